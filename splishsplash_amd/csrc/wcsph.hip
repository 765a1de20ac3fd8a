// wcsph.hip -- the WCSPH time step (SPlisHSPlasH/WCSPH/TimeStepWCSPH.cpp) on
// the DFSPH handle's device state: the shared neighbour lists, boundary
// records, density sweep and non-pressure forces, then the state equation
// p = k ((rho / rho0)^gamma - 1) on the clamped density, the pressure
// accelerations and the integration with the step's starting h.
#include <cmath>

#include "dfsph.hpp"

namespace sph {

namespace {

// density = max(density, rho0); p = k ((rho / rho0)^gamma - 1)
// (TimeStepWCSPH.cpp:98-108).
__global__ __launch_bounds__(BLOCK) void k_wc_pressure(Dev D, float k, float gamma) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float rho = fmaxf(D.rho[i], D.rho0);
    D.rho[i] = rho;
    D.p[i] = k * (powf(rho / D.rho0, gamma) - 1.0f);
}

// computePressureAccels (TimeStepWCSPH.cpp:137-199):
// a_i = -rho0 sum V (p_i / rho_i^2 + p_j / rho_j^2) grad W_ij and the boundary
// terms with p_i / rho_i^2 + p_i / rho0^2.
__global__ __launch_bounds__(BLOCK) void k_wc_accel(Dev D) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float4 xi = D.x[i];
    const float rhoi = D.rho[i], dpi = D.p[i] / (rhoi * rhoi);
    float ax = 0.f, ay = 0.f, az = 0.f;
    const unsigned cnt = D.cf[i];
    const unsigned *__restrict__ L = D.nf + i;
    const size_t S = D.stride;
    for (unsigned k = 0; k < cnt; k++) {
        const unsigned j = L[k * S];
        const float rhoj = D.rho[j];
        float3 r;
        const float g = kernel_G(xi, D.x[j], D.K, r);
        const float c = D.rho0 * D.V * (dpi + D.p[j] / (rhoj * rhoj));
        ax -= c * (g * r.x);
        ay -= c * (g * r.y);
        az -= c * (g * r.z);
    }
    const float dpb = dpi + D.p[i] / (D.rho0 * D.rho0);
    auto bterm = [&](const float4 &xb) {
        float3 r;
        const float g = kernel_G(xi, xb, D.K, r);
        const float c = D.rho0 * xb.w * dpb;
        ax -= c * (g * r.x);
        ay -= c * (g * r.y);
        az -= c * (g * r.z);
    };
    if (D.bmode == 1) {
        const unsigned cb = D.cb[i];
        for (unsigned k = 0; k < cb; k++) bterm(D.bx[D.nb[(size_t)k * D.stride + i]]);
    } else if (D.bmode == 2) {
        for (unsigned o = 0; o < D.nvm; o++) {
            const float4 xb = D.vmb[(size_t)o * D.stride + i];
            if (!(xb.w > 0.f)) break;
            bterm(xb);
        }
    } else if (D.bmode == 3) {
        for (unsigned o = 0; o < D.nvm; o++) {
            if (!(D.vmb[(size_t)o * D.stride + i].w > 0.f)) break;
            const float4 gb = D.dmg[(size_t)o * D.stride + i];
            ax += D.rho0 * dpb * gb.x;
            ay += D.rho0 * dpb * gb.y;
            az += D.rho0 * dpb * gb.z;
        }
    }
    D.ap[i] = make_float4(ax, ay, az, 0.f);
}

// a += a_p, v += h a, x += h v for active particles with the step's
// starting h (TimeStepWCSPH.cpp:111-128), then the time.
__global__ __launch_bounds__(BLOCK) void k_wc_final(Dev D) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    Ctl *C = D.ctl;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        C->t_prev = C->time;
        C->time += C->h_step;
        C->it = C->it_v = 0;  // no solver: the counters read 0
        C->avg_err = C->avg_err_v = 0.0;
    }
    if (i >= D.n) return;
    const bool act = D.state == nullptr || D.state[i] == 0u;
    if (!act) return;
    const float hs = (float)C->h_step;
    const float4 ap = D.ap[i];
    float4 a = D.a[i];
    a.x += ap.x;
    a.y += ap.y;
    a.z += ap.z;
    D.a[i] = a;
    float4 v = D.v[i];
    v.x += a.x * hs;
    v.y += a.y * hs;
    v.z += a.z * hs;
    D.v[i] = v;
    float4 x = D.x[i], c = D.xc[i];
    const float yx = hs * v.x - c.x, yy = hs * v.y - c.y, yz = hs * v.z - c.z;
    const float tx = x.x + yx, ty = x.y + yy, tz = x.z + yz;
    c.x = (tx - x.x) - yx;
    c.y = (ty - x.y) - yy;
    c.z = (tz - x.z) - yz;
    x.x = tx;
    x.y = ty;
    x.z = tz;
    D.x[i] = x;
    D.xc[i] = c;
}

}  // namespace

void DFSPH::set_wcsph(double stiffness, double exponent) {
    // both parameters have the minimum 1e-6 (TimeStepWCSPH.cpp:56-63)
    SPH_REQUIRE(std::isfinite(stiffness) && stiffness >= 1e-6 && std::isfinite(exponent) && exponent >= 1e-6,
                SPH_ERR_INVALID, "WCSPH: stiffness and exponent must be >= 1e-6");
    wc_k_ = stiffness;
    wc_gamma_ = exponent;
}

// TimeStepWCSPH::step (TimeStepWCSPH.cpp:74-131).
void DFSPH::step_wcsph() {
    neighborhood_search();
    volume_maps();
    sweep_init(true, false, false);  // computeDensities
    const bool weiler = P.viscosity_method == SPH_VISCOSITY_WEILER2018;
    const bool stdv = P.viscosity_method == SPH_VISCOSITY_STANDARD;
    if ((weiler || stdv || P.xsph != 0.0) && n_) mrho_copy();
    if (weiler) viscosity(false);
    const int expl = ((P.xsph != 0.0 || P.xsph_boundary != 0.0) ? 1 : 0) | (stdv ? 2 : 0);
    // the CFL update follows the pressure in the reference; it reads the
    // non-pressure acceleration only, so it can share this sweep
    post_forces(true, weiler, expl, P.cfl_method == 1 || P.cfl_method == 2, false, false);
    mrho_w_ = false;
    const Dev D = dev();
    const unsigned nb = nblocks(n_);
    {
        TimedScope ts(timer, stream, "wcsph_pressure");
        k_wc_pressure<<<nb, BLOCK, 0, stream>>>(D, (float)wc_k_, (float)wc_gamma_);
    }
    {
        TimedScope ts(timer, stream, "wcsph_accel");
        k_wc_accel<<<nb, BLOCK, 0, stream>>>(D);
    }
    {
        TimedScope ts(timer, stream, "wcsph_final");
        k_wc_final<<<nb, BLOCK, 0, stream>>>(D);
    }
    SPH_HIP(hipGetLastError());
    itLast_ = itVLast_ = 0;
    emitter_system_step();
    steps_++;
}

}  // namespace sph
