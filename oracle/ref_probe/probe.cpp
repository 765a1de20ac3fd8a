// probe.cpp -- TEST INFRASTRUCTURE. Links the reference's own
// SPlisHSPlasH/SPHKernels.cpp and SPlisHSPlasH/Utilities/GaussQuadrature.cpp
// (FP64 build, -DUSE_DOUBLE) and the header-only Eigen CG with the
// reference's BlockJacobiPreconditioner3D (Utilities/MatrixFreeSolver.h),
// and prints golden vectors as JSON on stdout. tests/golden/make_golden.py
// runs it and stores the output as fixtures.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>

#include "SPlisHSPlasH/Common.h"
#include "SPlisHSPlasH/SPHKernels.h"
#include "SPlisHSPlasH/Utilities/GaussQuadrature.h"
#include "SPlisHSPlasH/Utilities/MatrixFreeSolver.h"
#include <Eigen/IterativeLinearSolvers>

namespace SPH {
extern unsigned int const gaussian_n_1[101];
extern double const gaussian_abscissae_1[101][51];
extern double const gaussian_weights_1[101][51];
}  // namespace SPH

using namespace SPH;

// Access to the protected tables of the precomputed kernel.
struct Peek : public PrecomputedKernel<CubicKernel, 10000> {
    static double T(unsigned i) { return m_W[i]; }
    static double G(unsigned i) { return m_gradW[i]; }
    static double invStep() { return m_invStepSize; }
};

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double urand() {  // xorshift64*, fixed seed, in [0,1)
    g_state ^= g_state >> 12;
    g_state ^= g_state << 25;
    g_state ^= g_state >> 27;
    return (double)((g_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

static void arr(const char *name, const std::vector<double> &v, bool last = false) {
    printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); i++) printf("%s%.17g", i ? ", " : "", v[i]);
    printf("]%s\n", last ? "" : ",");
}

// Dense matrix for the CG probe
static std::vector<double> g_A;
static size_t g_n = 0;
static void matvec(const Real *vec, Real *res, void *) {
    for (size_t r = 0; r < g_n; r++) {
        double s = 0.0;
        for (size_t c = 0; c < g_n; c++) s += g_A[r * g_n + c] * vec[c];
        res[r] = s;
    }
}
static void diag(const unsigned int i, Matrix3r &m, void *) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) m(r, c) = g_A[(3 * i + r) * g_n + 3 * i + c];
}

int main() {
    printf("{\n");
    // --- precomputed cubic kernel tables and samples, three support radii
    const double radii[3] = {0.1, 0.0375, 0.0128};
    printf("\"kernels\": [\n");
    for (int ri = 0; ri < 3; ri++) {
        const double h = radii[ri];
        PrecomputedKernel<CubicKernel, 10000>::setRadius(h);
        CubicKernel::setRadius(h);
        std::vector<double> idx, T, G, rs, W, gW, Wc, gWc;
        double sumT = 0, sumG = 0;
        for (unsigned i = 0; i < 10000; i++) sumT += Peek::T(i);
        for (unsigned i = 0; i < 10001; i++) sumG += Peek::G(i);
        for (unsigned i = 0; i < 10000; i += 37) {
            idx.push_back(i);
            T.push_back(Peek::T(i));
            G.push_back(Peek::G(i));
        }
        idx.push_back(9999);
        T.push_back(Peek::T(9999));
        G.push_back(Peek::G(9999));
        for (int s = 0; s < 256; s++) {
            // radius in [0, 1.05 h], random direction
            const double rr = 1.05 * h * urand();
            double d[3];
            double nn = 0;
            do {
                for (int c = 0; c < 3; c++) d[c] = 2 * urand() - 1;
                nn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            } while (nn < 1e-3 || nn > 1.0);
            Vector3r r(d[0] / nn * rr, d[1] / nn * rr, d[2] / nn * rr);
            for (int c = 0; c < 3; c++) rs.push_back(r[c]);
            W.push_back(PrecomputedKernel<CubicKernel, 10000>::W(r));
            Vector3r g = PrecomputedKernel<CubicKernel, 10000>::gradW(r);
            for (int c = 0; c < 3; c++) gW.push_back(g[c]);
            Wc.push_back(CubicKernel::W(r));
            Vector3r gc = CubicKernel::gradW(r);
            for (int c = 0; c < 3; c++) gWc.push_back(gc[c]);
        }
        printf("{\"h\": %.17g, \"W_zero\": %.17g, \"invStep\": %.17g, \"sumT\": %.17g, \"sumG\": %.17g,\n", h,
               PrecomputedKernel<CubicKernel, 10000>::W_zero(), Peek::invStep(), sumT, sumG);
        arr("index", idx);
        arr("T", T);
        arr("G", G);
        arr("r", rs);
        arr("W", W);
        arr("gradW", gW);
        arr("W_cubic", Wc);
        arr("gradW_cubic", gWc, true);
        printf("}%s\n", ri < 2 ? "," : "");
    }
    printf("],\n");

    // --- Gauss-Legendre rule used by the volume map (p = 30)
    {
        const unsigned p = 30, n = gaussian_n_1[p];
        std::vector<double> x, w;
        for (unsigned i = 0; i < n; i++) {
            x.push_back(gaussian_abscissae_1[p][i]);
            w.push_back(gaussian_weights_1[p][i]);
        }
        printf("\"gauss_p30\": {\"n\": %u,\n", n);
        arr("x", x);
        arr("w", w, true);
        printf("},\n");
    }
    // --- GaussQuadrature::integrate on monomials over a box
    {
        printf("\"gauss_monomials\": [\n");
        const int exps[6][3] = {{0, 0, 0}, {2, 0, 0}, {2, 2, 2}, {4, 1, 0}, {6, 4, 2}, {10, 8, 6}};
        Eigen::AlignedBox3d dom(Eigen::Vector3d(-0.1, -0.2, 0.05), Eigen::Vector3d(0.3, 0.1, 0.25));
        for (int e = 0; e < 6; e++) {
            auto f = [&](Eigen::Vector3d const &x) {
                return std::pow(x[0], exps[e][0]) * std::pow(x[1], exps[e][1]) * std::pow(x[2], exps[e][2]);
            };
            const double v = GaussQuadrature::integrate(f, dom, 30);
            printf("{\"exp\": [%d, %d, %d], \"lo\": [-0.1, -0.2, 0.05], \"hi\": [0.3, 0.1, 0.25], \"value\": %.17g}%s\n",
                   exps[e][0], exps[e][1], exps[e][2], v, e < 5 ? "," : "");
        }
        printf("],\n");
    }
    // --- volume integrand of SimulatorBase.cpp:2637-2650 for a half-space
    // wall (sdf(y) = y_1 - wall), i.e. the volume of a node at distance d
    {
        const double h = 0.1;
        CubicKernel::setRadius(h);
        auto int_domain = Eigen::AlignedBox3d(Eigen::Vector3d::Constant(-h), Eigen::Vector3d::Constant(h));
        printf("\"halfspace_volume\": {\"h\": 0.1, ");
        std::vector<double> ds, vs;
        for (int s = -12; s <= 24; s++) {
            const double d = s * (h / 10.0);
            auto integrand = [&](Eigen::Vector3d const &xi) -> double {
                if (xi.squaredNorm() > h * h) return 0.0;
                const double dist = d + xi[1];
                if (dist <= 0.0) return 1.0;
                if (dist < h) return static_cast<double>(CubicKernel::W(static_cast<Real>(dist)) / CubicKernel::W_zero());
                return 0.0;
            };
            ds.push_back(d);
            vs.push_back(d > 2.0 * h ? 0.0 : 0.8 * GaussQuadrature::integrate(integrand, int_domain, 30));
        }
        arr("d", ds);
        arr("volume", vs, true);
        printf("},\n");
    }
    // --- Eigen CG with BlockJacobiPreconditioner3D on an SPD system
    {
        const size_t m = 12;
        g_n = 3 * m;
        std::vector<double> B(g_n * g_n);
        for (auto &b : B) b = urand() - 0.5;
        g_A.assign(g_n * g_n, 0.0);
        for (size_t r = 0; r < g_n; r++)
            for (size_t c = 0; c < g_n; c++) {
                double s = 0;
                for (size_t k = 0; k < g_n; k++) s += B[r * g_n + k] * B[c * g_n + k];
                g_A[r * g_n + c] = s + (r == c ? 0.5 : 0.0);
            }
        VectorXr b(g_n), x0(g_n);
        for (size_t i = 0; i < g_n; i++) {
            b[i] = urand() - 0.5;
            x0[i] = 0.1 * (urand() - 0.5);
        }
        std::vector<double> bv(b.data(), b.data() + g_n), x0v(x0.data(), x0.data() + g_n);
        printf("\"cg\": [\n");
        const double tols[2] = {1e-4, 1e-10};
        for (int t = 0; t < 2; t++) {
            MatrixReplacement A((unsigned)g_n, matvec, nullptr);
            Eigen::ConjugateGradient<MatrixReplacement, Eigen::Lower | Eigen::Upper, BlockJacobiPreconditioner3D> solver;
            solver.preconditioner().init((unsigned)m, diag, nullptr);
            solver.setTolerance(tols[t]);
            solver.setMaxIterations(100);
            solver.compute(A);
            VectorXr x = solver.solveWithGuess(b, x0);
            std::vector<double> xv(x.data(), x.data() + g_n);
            printf("{\"n\": %zu, \"tol\": %.17g, \"maxIter\": 100, \"iterations\": %d, \"error\": %.17g,\n", g_n, tols[t],
                   (int)solver.iterations(), (double)solver.error());
            arr("A", g_A);
            arr("b", bv);
            arr("x0", x0v);
            arr("x", xv, true);
            printf("}%s\n", t == 0 ? "," : "");
        }
        printf("]\n");
    }
    printf("}\n");
    return 0;
}
