// Drives the C++ facade (include/drifter_hip/primitive_equations_gpu.hpp) the
// way the reference's SimulationDriver drives its solvers, and checks every
// result against the FP64 oracle (oracle/liboracle.so).  Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "drifter_hip/primitive_equations_gpu.hpp"
#include "oracle.h"

using drifter_gpu::Vec;

static double rel_l2(const Vec &a, const Vec &b, size_t off, size_t stride, size_t len, size_t nblk)
{
    double num = 0, den = 0;
    for (size_t k = 0; k < nblk; ++k)
        for (size_t i = 0; i < len; ++i) {
            const double x = a[k * stride + off + i], y = b[k * stride + off + i];
            num += (x - y) * (x - y);
            den += y * y;
        }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

int main()
{
    const int order = 3, n = order + 1, nd = n * n * n, E = 24;
    double x[9], w[9];
    if (drf_basis(order, x, w, nullptr) != DRF_OK)
        return 2;
    // a smooth state on E elements of a 6 x 4 strip (each element 1 km wide)
    std::vector<Vec> h(E, Vec(nd)), y(E, Vec(nd)), zero(E, Vec(nd, 0.0));
    Vec U(E * 5 * nd), hflat(E * nd), fflat(E * nd);
    const drifter_gpu::CoriolisParameter cor{1e-4, 2e-11, 0.0};
    for (int e = 0; e < E; ++e)
        for (int k = 0; k < n; ++k)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) {
                    const int d = i + n * (j + n * k);
                    const double X = 1000.0 * (e % 6) + 500.0 * (1 + x[i]);
                    const double Y = 1000.0 * (e / 6) + 500.0 * (1 + x[j]);
                    const double S = 0.5 * (1 + x[k]);
                    const double hb = 40.0 + 5.0 * std::sin(X / 1500.0) * std::cos(Y / 2100.0);
                    const double eta = 0.05 * std::cos(X / 900.0 + Y / 1300.0);
                    const double u = 0.1 * std::sin(Y / 700.0 + S), v = 0.08 * std::cos(X / 800.0 - S);
                    const double T = 15 + 3 * std::sin(X / 1100.0 + 2 * S), Sal = 33 + std::cos(Y / 1700.0);
                    const double H = eta + hb;
                    h[e][d] = hb;
                    y[e][d] = Y;
                    hflat[e * nd + d] = hb;
                    fflat[e * nd + d] = cor(Y);
                    double *B = &U[(size_t)e * 5 * nd];
                    B[d] = H * u;
                    B[nd + d] = H * v;
                    B[2 * nd + d] = eta;
                    B[3 * nd + d] = H * T;
                    B[4 * nd + d] = H * Sal;
                }
    orc_physics ph;
    orc_default_physics(&ph);
    int fails = 0;
    try {
        drifter_gpu::PrimitiveEquationsSolver solver(order);
        try { // RHS before initialize is an error, not a crash
            Vec tmp;
            solver.compute_rhs(0.0, U, tmp);
            std::printf("FAIL: no error before initialize\n");
            ++fails;
        } catch (const drifter_gpu::Error &err) {
            if (err.status != DRF_ERR_STATE)
                ++fails;
        }
        solver.initialize(E, h, zero, zero, cor, y);
        drifter_gpu::RHSFunction rhs = solver.rhs_function();
        Vec dU, ref(U.size());
        rhs(0.0, U, dU);
        orc_rhs(order, E, U.data(), hflat.data(), fflat.data(), &ph, ref.data());
        for (int f = 0; f < 5; ++f) {
            const double r = rel_l2(dU, ref, f * nd, 5 * nd, nd, E);
            std::printf("rhs field %d rel L2 %.3e\n", f, r);
            fails += !(r <= 1e-12);
        }
        // one driver-like RK4 step: dt from the CFL controller
        const double c = solver.max_wave_speed(U);
        drifter_gpu::AdaptiveTimeController ctl(0.5, 1e-6, 3600.0);
        const double dt = ctl.compute_dt(c, 2.0 / n);
        auto stepper = drifter_gpu::TimeStepper::rk4_classic();
        Vec Ug = U, Uo = U;
        stepper->step(0.0, dt, Ug, solver);
        orc_rk_step(ORC_RK4, order, E, 0.0, dt, Uo.data(), hflat.data(), fflat.data(), &ph);
        for (int f = 0; f < 5; ++f) {
            const double r = rel_l2(Ug, Uo, f * nd, 5 * nd, nd, E);
            std::printf("rk4 field %d rel L2 %.3e (dt %.3e)\n", f, r, dt);
            fails += !(r <= 1e-12);
        }
        // mode split on the full state
        Vec full((size_t)E * 12 * nd, 0.0);
        for (int e = 0; e < E; ++e)
            for (int d = 0; d < nd; ++d) {
                const double *B = &U[(size_t)e * 5 * nd];
                double *F = &full[(size_t)e * 12 * nd];
                const double Hh = B[2 * nd + d] + hflat[e * nd + d];
                const double hi = Hh > 1e-10 ? 1.0 / Hh : 0.0;
                for (int q = 0; q < 5; ++q)
                    F[q * nd + d] = B[q * nd + d];
                F[5 * nd + d] = Hh;
                F[6 * nd + d] = B[d] * hi;
                F[7 * nd + d] = B[nd + d] * hi;
                F[8 * nd + d] = B[3 * nd + d] * hi;
                F[9 * nd + d] = B[4 * nd + d] * hi;
            }
        Vec fg = full, fo = full;
        drifter_gpu::ModeSplittingSolver ms(solver, 20);
        ms.step(dt, fg);
        orc_mode_split_step(order, E, dt, 20, fo.data(), hflat.data(), fflat.data(), &ph);
        for (int q = 0; q < 12; ++q) {
            const double r = rel_l2(fg, fo, q * nd, 12 * nd, nd, E);
            fails += !(r <= 1e-12);
            if (q < 6 || !(r <= 1e-12))
                std::printf("mode split field %d rel L2 %.3e\n", q, r);
        }
        // MortarSpace operators (host copies) against the oracle's
        for (int t = DRF_CONN_SAME_LEVEL; t <= DRF_CONN_FINE_2X2; ++t) {
            drifter_gpu::MortarSpace m(order, t, -1);
            const int nm = m.num_dofs(), nf = m.num_fine_elements();
            Vec M(nm * nm), Mi(nm * nm), Pc(nm * nm), Lc(nm * nm), Pf(4 * nm * nm),
                Lf(4 * nm * nm);
            const int onf = orc_mortar_operators(order, t, M.data(), Mi.data(), Pc.data(),
                                                 Lc.data(), Pf.data(), Lf.data());
            Vec gM = m.mass_matrix(), gPf = m.projection_fine(), gLf = m.lift_fine();
            double dM = 0, dP = 0, dL = 0;
            for (int i = 0; i < nm * nm; ++i)
                dM = std::fmax(dM, std::fabs(gM[i] - M[i]));
            for (int i = 0; i < nf * nm * nm; ++i) {
                dP = std::fmax(dP, std::fabs(gPf[i] - Pf[i]));
                dL = std::fmax(dL, std::fabs(gLf[i] - Lf[i]));
            }
            const bool ok = onf == nf && dM == 0.0 && dL == 0.0 && dP < 1e-12;
            fails += !ok;
            std::printf("mortar type %d: %d fine, |dM| %.1e |dPf| %.1e |dLf| %.1e %s\n", t, nf,
                        dM, dP, dL, ok ? "ok" : "MISMATCH");
        }
        // AdaptiveMeshRefinement::transfer_solution against the oracle
        {
            const int to = 3, tnd = 64, tE = 20;
            std::vector<Vec> old_sol(tE, Vec(tnd));
            Vec packed(tE * tnd);
            unsigned seed = 12345u;
            for (int e = 0; e < tE; ++e)
                for (int d = 0; d < tnd; ++d) {
                    seed = seed * 1664525u + 1013904223u;
                    old_sol[e][d] = packed[e * tnd + d] = (double)(seed >> 8) / 16777216.0 - 0.5;
                }
            std::vector<std::pair<int64_t, std::vector<int64_t>>> rmap;
            const int64_t par[3] = {2, 5, 19};
            for (int64_t p : par)
                rmap.push_back({p, std::vector<int64_t>(8, -1)});
            std::vector<Vec> got(1, Vec(tnd, 7.0)); // appended to, as the reference does
            drifter_gpu::transfer_solution(to, old_sol, rmap, got);
            Vec ref((tE + 21) * tnd);
            orc_transfer_solution(to, tE, packed.data(), 3, par, ref.data());
            double dmax = 0.0, scale = 0.0;
            bool copies_exact = got.size() == (size_t)(1 + tE + 21) && got[0][0] == 7.0;
            for (int k = 0; copies_exact && k < tE + 21; ++k)
                for (int d = 0; d < tnd; ++d) {
                    const double a = got[1 + k][d], b = ref[(size_t)k * tnd + d];
                    if (k < tE - 3 && a != b)
                        copies_exact = false;
                    dmax = std::fmax(dmax, std::fabs(a - b));
                    scale = std::fmax(scale, std::fabs(b));
                }
            const bool ok = copies_exact && dmax <= 1e-13 * scale;
            fails += !ok;
            std::printf("transfer_solution: %zu elements, max |diff| %.1e %s\n", got.size() - 1,
                        dmax, ok ? "ok" : "MISMATCH");
        }
    } catch (const std::exception &ex) {
        std::printf("FAIL: exception %s\n", ex.what());
        return 1;
    }
    std::printf(fails ? "FAILED (%d)\n" : "PASSED\n", fails);
    return fails ? 1 : 0;
}
