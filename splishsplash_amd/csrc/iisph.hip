// iisph.hip -- the IISPH time step (SPlisHSPlasH/IISPH/TimeStepIISPH.cpp) on
// the DFSPH handle's device state: the same neighbour lists, boundary
// records, density sweep and non-pressure forces, with IISPH's relaxed
// Jacobi solve (omega = 0.5) in place of the two DFSPH solves.
//
// predictAdvection is two sweeps (d_ii with v += h a; rho_adv, a_ii and the
// half pressure warm start), an iteration three (sum_j d_ij p_j; the
// pressure update with the FP64 error sum; the new pressure becomes the
// last one) and a one-block reduction whose decision takes effect after the
// iteration, so launches past convergence are no-ops.
#include <algorithm>
#include <cmath>

#include "dfsph.hpp"

namespace sph {

namespace {

__device__ __forceinline__ bool active_ii(const Dev &D, unsigned i) { return D.state == nullptr || D.state[i] == 0u; }

struct IIArgs {
    float4 *dii, *dijpj;  // d_ii, sum_j d_ij p_j
    float *lastp, *q;     // last pressure, last pressure / rho^2
};

// d_ii = -sum V / rho_i^2 grad W_ij and the boundary terms; v += h a for
// active particles (TimeStepIISPH.cpp:143-196).
__global__ __launch_bounds__(BLOCK) void k_ii_dii(Dev D, IIArgs A) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float h = (float)D.ctl->h;
    if (active_ii(D, i)) {
        const float4 a = D.a[i];
        float4 v = D.v[i];
        v.x += h * a.x;
        v.y += h * a.y;
        v.z += h * a.z;
        D.v[i] = v;
    }
    const float4 xi = D.x[i];
    const float rho = D.rho[i] / D.rho0, inv2 = 1.0f / (rho * rho);
    float dx = 0.f, dy = 0.f, dz = 0.f;
    const unsigned cnt = D.cf[i];
    const unsigned *__restrict__ L = D.nf + i;
    const size_t S = D.stride;
    for (unsigned k = 0; k < cnt; k++) {
        float3 r;
        const float g = kernel_G(xi, D.x[L[k * S]], D.K, r);
        const float c = D.V * inv2;
        dx -= c * (g * r.x);
        dy -= c * (g * r.y);
        dz -= c * (g * r.z);
    }
    if (D.bmode == 1) {
        const unsigned cb = D.cb[i];
        for (unsigned k = 0; k < cb; k++) {
            const float4 xb = D.bx[D.nb[(size_t)k * D.stride + i]];
            float3 r;
            const float g = kernel_G(xi, xb, D.K, r);
            const float c = xb.w * inv2;
            dx -= c * (g * r.x);
            dy -= c * (g * r.y);
            dz -= c * (g * r.z);
        }
    } else if (D.bmode == 2) {
        for (unsigned o = 0; o < D.nvm; o++) {
            const float4 xb = D.vmb[(size_t)o * D.stride + i];
            if (!(xb.w > 0.f)) break;
            float3 r;
            const float g = kernel_G(xi, xb, D.K, r);
            const float c = xb.w * inv2;
            dx -= c * (g * r.x);
            dy -= c * (g * r.y);
            dz -= c * (g * r.z);
        }
    } else if (D.bmode == 3) {
        for (unsigned o = 0; o < D.nvm; o++) {
            if (!(D.vmb[(size_t)o * D.stride + i].w > 0.f)) break;
            const float4 gb = D.dmg[(size_t)o * D.stride + i];
            dx += inv2 * gb.x;
            dy += inv2 * gb.y;
            dz += inv2 * gb.z;
        }
    }
    A.dii[i] = make_float4(dx, dy, dz, 0.f);
}

// rho_adv = rho_i + h sum V (v_i - v_j) . grad W_ij (+ boundary), the last
// pressure p / 2 and a_ii = sum V (d_ii - d_ji) . grad W_ij (+ boundary),
// d_ji = V / rho_i^2 grad W_ij (TimeStepIISPH.cpp:198-287).
__global__ __launch_bounds__(BLOCK) void k_ii_adv(Dev D, IIArgs A) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float h = (float)D.ctl->h;
    const float4 xi = D.x[i], vi = D.v[i], dii = A.dii[i];
    const float rho = D.rho[i] / D.rho0, rho2 = rho * rho, dpi = D.V / rho2;
    float adv = 0.f, aii = 0.f;
    const unsigned cnt = D.cf[i];
    const unsigned *__restrict__ L = D.nf + i;
    const size_t S = D.stride;
    for (unsigned k = 0; k < cnt; k++) {
        const unsigned j = L[k * S];
        const float4 vj = D.v[j];
        float3 r;
        const float g = kernel_G(xi, D.x[j], D.K, r);
        const float gx = g * r.x, gy = g * r.y, gz = g * r.z;
        adv += h * D.V * (((vi.x - vj.x) * gx + (vi.y - vj.y) * gy) + (vi.z - vj.z) * gz);
        aii += D.V * (((dii.x - dpi * gx) * gx + (dii.y - dpi * gy) * gy) + (dii.z - dpi * gz) * gz);
    }
    auto bpair = [&](const float4 &xb) {
        float3 r;
        const float g = kernel_G(xi, xb, D.K, r);
        const float gx = g * r.x, gy = g * r.y, gz = g * r.z;
        adv += h * xb.w * ((vi.x * gx + vi.y * gy) + vi.z * gz);
        aii += xb.w * (((dii.x - dpi * gx) * gx + (dii.y - dpi * gy) * gy) + (dii.z - dpi * gz) * gz);
    };
    if (D.bmode == 1) {
        const unsigned cb = D.cb[i];
        for (unsigned k = 0; k < cb; k++) bpair(D.bx[D.nb[(size_t)k * D.stride + i]]);
    } else if (D.bmode == 2) {
        for (unsigned o = 0; o < D.nvm; o++) {
            const float4 xb = D.vmb[(size_t)o * D.stride + i];
            if (!(xb.w > 0.f)) break;
            bpair(xb);
        }
    } else if (D.bmode == 3) {
        for (unsigned o = 0; o < D.nvm; o++) {
            if (!(D.vmb[(size_t)o * D.stride + i].w > 0.f)) break;
            const float4 gb = D.dmg[(size_t)o * D.stride + i];
            adv -= h * ((vi.x * gb.x + vi.y * gb.y) + vi.z * gb.z);
            const float jx = -gb.x / rho2, jy = -gb.y / rho2, jz = -gb.z / rho2;  // d_ji
            aii -= ((dii.x - jx) * gb.x + (dii.y - jy) * gb.y) + (dii.z - jz) * gb.z;
        }
    }
    D.radv[i] = rho + adv;
    const float lp = 0.5f * D.p[i];
    A.lastp[i] = lp;
    A.q[i] = lp / rho2;
    D.alpha[i] = aii;  // a_ii
}

// sum_j d_ij p_j = -sum V p_j / rho_j^2 grad W_ij (TimeStepIISPH.cpp:355-377).
__global__ __launch_bounds__(BLOCK) void k_ii_dijpj(Dev D, IIArgs A) {
    if (D.ctl->done) return;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float4 xi = D.x[i];
    float sx = 0.f, sy = 0.f, sz = 0.f;
    const unsigned cnt = D.cf[i];
    const unsigned *__restrict__ L = D.nf + i;
    const size_t S = D.stride;
    for (unsigned k = 0; k < cnt; k++) {
        const unsigned j = L[k * S];
        float3 r;
        const float g = kernel_G(xi, D.x[j], D.K, r);
        const float c = D.V * A.q[j];
        sx -= c * (g * r.x);
        sy -= c * (g * r.y);
        sz -= c * (g * r.z);
    }
    A.dijpj[i] = make_float4(sx, sy, sz, 0.f);
}

// The relaxed Jacobi update and the error of the density it predicts
// (TimeStepIISPH.cpp:379-454); per-block FP64 sums.
__global__ __launch_bounds__(BLOCK) void k_ii_press(Dev D, IIArgs A) {
    __shared__ double lds[BLOCK / 64];
    Ctl *C = D.ctl;
    if (C->done) return;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    double err = 0.0;
    if (i < D.n) {
        const double hd = C->h;
        const float h2 = (float)(hd * hd);
        const float4 xi = D.x[i], dpj_i = A.dijpj[i];
        const float rho = D.rho[i] / D.rho0, dpi = D.V / (rho * rho), lpi = A.lastp[i];
        float sum = 0.f;
        const unsigned cnt = D.cf[i];
        const unsigned *__restrict__ L = D.nf + i;
        const size_t S = D.stride;
        for (unsigned k = 0; k < cnt; k++) {
            const unsigned j = L[k * S];
            float3 r;
            const float g = kernel_G(xi, D.x[j], D.K, r);
            const float gx = g * r.x, gy = g * r.y, gz = g * r.z;
            const float4 djj = A.dii[j], djk = A.dijpj[j];
            const float lpj = A.lastp[j];
            // d_ij p_j (i) - d_jj p_j - (sum_k d_jk p_k - d_ji p_i)
            const float ux = dpj_i.x - djj.x * lpj - (djk.x - dpi * gx * lpi);
            const float uy = dpj_i.y - djj.y * lpj - (djk.y - dpi * gy * lpi);
            const float uz = dpj_i.z - djj.z * lpj - (djk.z - dpi * gz * lpi);
            sum += D.V * ((ux * gx + uy * gy) + uz * gz);
        }
        auto bsum = [&](const float4 &xb) {
            float3 r;
            const float g = kernel_G(xi, xb, D.K, r);
            sum += xb.w * (((dpj_i.x * (g * r.x)) + dpj_i.y * (g * r.y)) + dpj_i.z * (g * r.z));
        };
        if (D.bmode == 1) {
            const unsigned cb = D.cb[i];
            for (unsigned k = 0; k < cb; k++) bsum(D.bx[D.nb[(size_t)k * D.stride + i]]);
        } else if (D.bmode == 2) {
            for (unsigned o = 0; o < D.nvm; o++) {
                const float4 xb = D.vmb[(size_t)o * D.stride + i];
                if (!(xb.w > 0.f)) break;
                bsum(xb);
            }
        } else if (D.bmode == 3) {
            for (unsigned o = 0; o < D.nvm; o++) {
                if (!(D.vmb[(size_t)o * D.stride + i].w > 0.f)) break;
                const float4 gb = D.dmg[(size_t)o * D.stride + i];
                sum -= (dpj_i.x * gb.x + dpj_i.y * gb.y) + dpj_i.z * gb.z;
            }
        }
        const float aii = D.alpha[i];
        const float b = 1.0f - D.radv[i];
        const float denom = aii * h2;
        float pi = 0.f;
        if (fabsf(denom) > 1.0e-9f) pi = fmaxf(0.5f * lpi + 0.5f / denom * (b - h2 * sum), 0.0f);
        D.p[i] = pi;
        if (pi != 0.0f) {
            const double rho0 = (double)D.rho0;
            const double nd = rho0 * (((double)aii * pi + (double)sum) * (double)h2 - (double)b) + rho0;
            err = nd - rho0;
        }
    }
    const double bs = block_sum(err, lds);
    if (threadIdx.x == 0) D.partial[blockIdx.x] = bs;
}

__global__ __launch_bounds__(RBLOCK) void k_ii_reduce(Dev D, unsigned nparts) {
    __shared__ double lds[RBLOCK / 64];
    Ctl *C = D.ctl;
    if (C->done) return;
    double t = 0.0;
    for (unsigned b = threadIdx.x; b < nparts; b += blockDim.x) t += D.partial[b];
    t = block_sum(t, lds);
    if (threadIdx.x == 0) {
        const double avg = D.n ? t / (double)D.n : 0.0;
        const bool finite = isfinite(avg);
        if (!finite) C->flags |= 1;
        C->avg_err = avg;
        C->it += 1;
        const double eta = D.maxError * 0.01 * (double)D.rho0;
        C->done_v = ((avg <= eta && C->it >= D.minIt) || C->it >= D.maxIt || !finite) ? 1 : 0;  // pending
    }
}

// lastPressure = pressure (TimeStepIISPH.cpp:456-461).
__global__ __launch_bounds__(BLOCK) void k_ii_last(Dev D, IIArgs A) {
    if (D.ctl->done) return;
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float p = D.p[i], rho = D.rho[i] / D.rho0;
    A.lastp[i] = p;
    A.q[i] = p / (rho * rho);
}

__global__ void k_ii_commit(Ctl *C) {
    if (!C->done) C->done = C->done_v;
}

__global__ void k_ii_reset(Ctl *C) {
    C->it = 0;
    C->done = 0;
    C->done_v = 0;
    C->avg_err = 0.0;
}

// computePressureAccels (TimeStepIISPH.cpp:489-554) and the integration
// with the current h (:464-487): v += h a_p, x += h v for active particles.
__global__ __launch_bounds__(BLOCK) void k_ii_accel(Dev D, IIArgs A) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const float4 xi = D.x[i];
    const float rho = D.rho[i] / D.rho0, dpi = D.p[i] / (rho * rho);
    float ax = 0.f, ay = 0.f, az = 0.f;
    const unsigned cnt = D.cf[i];
    const unsigned *__restrict__ L = D.nf + i;
    const size_t S = D.stride;
    for (unsigned k = 0; k < cnt; k++) {
        const unsigned j = L[k * S];
        float3 r;
        const float g = kernel_G(xi, D.x[j], D.K, r);
        const float c = D.V * (dpi + A.q[j]);
        ax -= c * (g * r.x);
        ay -= c * (g * r.y);
        az -= c * (g * r.z);
    }
    auto bterm = [&](const float4 &xb) {
        float3 r;
        const float g = kernel_G(xi, xb, D.K, r);
        const float c = xb.w * dpi;
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
            ax += dpi * gb.x;
            ay += dpi * gb.y;
            az += dpi * gb.z;
        }
    }
    D.ap[i] = make_float4(ax, ay, az, 0.f);
}

__global__ __launch_bounds__(BLOCK) void k_ii_final(Dev D) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    Ctl *C = D.ctl;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        C->t_prev = C->time;
        C->time += C->h_step;  // the step's starting h (TimeStepIISPH.cpp:127)
    }
    if (i >= D.n || !active_ii(D, i)) return;
    const float h = (float)C->h;  // the integration's h: after the CFL update
    const float4 ap = D.ap[i];
    float4 v = D.v[i];
    v.x += h * ap.x;
    v.y += h * ap.y;
    v.z += h * ap.z;
    D.v[i] = v;
    float4 x = D.x[i], c = D.xc[i];
    const float yx = h * v.x - c.x, yy = h * v.y - c.y, yz = h * v.z - c.z;
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

// TimeStepIISPH::step (TimeStepIISPH.cpp:91-128).
void DFSPH::step_iisph() {
    neighborhood_search();
    volume_maps();
    sweep_init(true, false, false);  // computeDensities
    const bool weiler = P.viscosity_method == SPH_VISCOSITY_WEILER2018;
    const bool stdv = P.viscosity_method == SPH_VISCOSITY_STANDARD;
    if ((weiler || stdv || P.xsph != 0.0) && n_) mrho_copy();
    if (weiler) viscosity(false);
    const int expl = ((P.xsph != 0.0 || P.xsph_boundary != 0.0) ? 1 : 0) | (stdv ? 2 : 0);
    post_forces(true, weiler, expl, P.cfl_method == 1 || P.cfl_method == 2, false, false);
    mrho_w_ = false;
    const size_t S = stride_;
    iid_.reserve(S);
    iidp_.reserve(S);
    ilp_.reserve(S);
    iq_.reserve(S);
    IIArgs A{iid_.get(), iidp_.get(), ilp_.get(), iq_.get()};
    const Dev D = dev();
    const unsigned nb = nblocks(n_);
    auto launch = [&](const char *name, auto kern) {
        TimedScope ts(timer, stream, name);
        kern<<<nb, BLOCK, 0, stream>>>(D, A);
        SPH_HIP(hipGetLastError());
    };
    launch("iisph_dii", k_ii_dii);
    launch("iisph_advection", k_ii_adv);
    k_ii_reset<<<1, 1, 0, stream>>>(ctl_.get());
    int launched = 0, chunk = std::max(P.min_iterations, std::max(1, lastIt_));
    for (;;) {
        for (int c = 0; c < chunk && launched < P.max_iterations; c++, launched++) {
            launch("iisph_dijpj", k_ii_dijpj);
            launch("iisph_pressure", k_ii_press);
            {
                TimedScope ts(timer, stream, "iisph_reduce");
                k_ii_reduce<<<1, RBLOCK, 0, stream>>>(D, nb);
            }
            launch("iisph_last", k_ii_last);
            k_ii_commit<<<1, 1, 0, stream>>>(ctl_.get());
            SPH_HIP(hipGetLastError());
        }
        read_ctl();
        if (hctl_->done || launched >= P.max_iterations) break;
        chunk = 2;
    }
    itLast_ = hctl_->it;
    lastIt_ = std::max(P.min_iterations, itLast_);
    totIt_ += itLast_;
    itVLast_ = 0;
    // the accelerations read p / rho^2 of the neighbours: the last pressure
    // is the final one (the iteration copied it)
    launch("iisph_accel", k_ii_accel);
    {
        TimedScope ts(timer, stream, "iisph_final");
        k_ii_final<<<nb, BLOCK, 0, stream>>>(D);
    }
    SPH_HIP(hipGetLastError());
    emitter_system_step();
    steps_++;
}

}  // namespace sph
