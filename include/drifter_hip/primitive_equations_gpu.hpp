// Header-only C++ facade over the drf C ABI, shaped like the reference's
// solver interfaces so that a C++ time integrator or mode splitter can call
// the MI355X right-hand side in place of the CPU operators:
//   PrimitiveEquationsSolver  include/physics/primitive_equations.hpp:217-265
//   RHSFunction / TimeStepper include/solver/time_stepper.hpp:10-33
//   ModeSplittingSolver       include/physics/mode_splitting.hpp:203-266
//   AdaptiveTimeController    include/solver/time_stepper.hpp:64-81
//   MortarSpace               include/dg/mortar.hpp:32-136
//   ConformingInterface       include/dg/mortar.hpp:187-215
//   GradientErrorEstimator    include/amr/refinement.hpp, src/amr/refinement.cpp:39-161
//   transfer_solution         AdaptiveMeshRefinement::transfer_solution, src/amr/refinement.cpp:533-567
// Vectors are std::vector<double> (Eigen's VecX maps onto them through
// .data()); failures throw std::runtime_error carrying drf_last_error(), as
// the reference reports errors by exception.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "drf.h"

namespace drifter_gpu {

using Real = double;
using Vec = std::vector<double>;
// RHSFunction, include/solver/time_stepper.hpp:10
using RHSFunction = std::function<void(Real t, const Vec &U, Vec &dUdt)>;

class Error : public std::runtime_error {
  public:
    Error(drf_status s, const std::string &what)
        : std::runtime_error("drf status " + std::to_string(s) + ": " + what), status(s)
    {
    }
    drf_status status;
};

inline void check(drf_status s)
{
    if (s != DRF_OK)
        throw Error(s, drf_last_error());
}

// CoriolisParameter, include/physics/primitive_equations.hpp:52-79
struct CoriolisParameter {
    Real f0 = 1e-4, beta = 2e-11, y0 = 0.0;
    Real operator()(Real y) const { return f0 + beta * (y - y0); }
};

class PrimitiveEquationsSolver {
  public:
    explicit PrimitiveEquationsSolver(int order, int precision = DRF_PREC_FP64, int device = 0)
        : order_(order)
    {
        drf_ctx *c = nullptr;
        check(drf_create(order, precision, device, &c));
        ctx_.reset(c);
    }

    int order() const { return order_; }
    int num_dofs() const { return (order_ + 1) * (order_ + 1) * (order_ + 1); }
    drf_ctx *ctx() const { return ctx_.get(); }

    void set_physics(const drf_physics &p) { check(drf_set_physics(ctx(), &p)); }

    // PrimitiveEquationsSolver::initialize (:229-231): per-element vectors
    void initialize(int num_elements, const std::vector<Vec> &bathymetry,
                    const std::vector<Vec> &dh_dx, const std::vector<Vec> &dh_dy,
                    const CoriolisParameter &coriolis, const std::vector<Vec> &y_positions)
    {
        (void)dh_dx;
        (void)dh_dy;
        const int nd = num_dofs();
        Vec h, y;
        h.reserve((size_t)num_elements * nd);
        for (int e = 0; e < num_elements; ++e) {
            if ((int)bathymetry.at(e).size() != nd)
                throw std::invalid_argument("bathymetry must hold ndof values per element");
            h.insert(h.end(), bathymetry[e].begin(), bathymetry[e].end());
        }
        if (!y_positions.empty())
            for (int e = 0; e < num_elements; ++e)
                y.insert(y.end(), y_positions.at(e).begin(), y_positions.at(e).end());
        check(drf_initialize(ctx(), num_elements, h.data(), nullptr, nullptr, coriolis.f0,
                             coriolis.beta, coriolis.y0, y.empty() ? nullptr : y.data()));
        num_elements_ = num_elements;
    }

    // SimulationDriver::compute_rhs on the packed vector (simulation_driver.cpp:170-197)
    void compute_rhs(Real t, const Vec &U, Vec &dUdt) const
    {
        dUdt.resize(U.size());
        check(drf_rhs(ctx(), t, U.data(), dUdt.data()));
    }

    RHSFunction rhs_function() const
    {
        return [this](Real t, const Vec &U, Vec &dUdt) { compute_rhs(t, U, dUdt); };
    }

    Real max_wave_speed(const Vec &U) const
    {
        Real c = 0.0;
        check(drf_max_wave_speed(ctx(), U.data(), &c));
        return c;
    }

    int num_elements() const { return num_elements_; }

  private:
    struct Del {
        void operator()(drf_ctx *c) const { drf_destroy(c); }
    };
    int order_;
    int num_elements_ = 0;
    std::unique_ptr<drf_ctx, Del> ctx_;
};

// TimeStepper with the reference's factories; step() runs fused on the GPU
// through the solver that owns the right-hand side.
class TimeStepper {
  public:
    static std::unique_ptr<TimeStepper> forward_euler() { return make(DRF_RK_EULER, 1, 1); }
    static std::unique_ptr<TimeStepper> rk2_midpoint() { return make(DRF_RK_RK2, 2, 2); }
    static std::unique_ptr<TimeStepper> rk3_ssp() { return make(DRF_RK_SSPRK3, 3, 3); }
    static std::unique_ptr<TimeStepper> rk4_classic() { return make(DRF_RK_RK4, 4, 4); }
    static std::unique_ptr<TimeStepper> lsrk45() { return make(DRF_RK_LSRK45, 4, 5); }
    static std::unique_ptr<TimeStepper> rk45_dormand_prince() { return make(DRF_RK_DP45, 5, 6); }
    // RKDG order selection, src/solver/time_stepper.cpp:209-220
    static std::unique_ptr<TimeStepper> rkdg(int order)
    {
        if (order <= 1)
            return forward_euler();
        if (order == 2)
            return rk2_midpoint();
        if (order == 3)
            return rk3_ssp();
        return rk4_classic();
    }

    void step(Real t, Real dt, Vec &U, const PrimitiveEquationsSolver &solver) const
    {
        check(drf_rk_step(solver.ctx(), scheme_, t, dt, U.data()));
    }
    int order() const { return order_; }
    int num_stages() const { return stages_; }
    int scheme() const { return scheme_; }

  private:
    TimeStepper(int scheme, int order, int stages) : scheme_(scheme), order_(order), stages_(stages)
    {
    }
    static std::unique_ptr<TimeStepper> make(int s, int o, int st)
    {
        return std::unique_ptr<TimeStepper>(new TimeStepper(s, o, st));
    }
    int scheme_, order_, stages_;
};

// AdaptiveTimeController::compute_dt (src/solver/time_stepper.cpp:261-268)
class AdaptiveTimeController {
  public:
    explicit AdaptiveTimeController(Real cfl = 0.5, Real dt_min = 1.0e-10, Real dt_max = 1.0e6)
        : cfl_(cfl), dt_min_(dt_min), dt_max_(dt_max)
    {
    }
    Real compute_dt(Real max_wave_speed, Real min_element_size) const
    {
        return drf_compute_dt(max_wave_speed, min_element_size, cfl_, dt_min_, dt_max_);
    }

  private:
    Real cfl_, dt_min_, dt_max_;
};

// ModeSplittingSolver::step on the flat full state
// [elem][Hu|Hv|eta|HT|HS|H|u|v|T|S|omega|rho][ndof].
class ModeSplittingSolver {
  public:
    ModeSplittingSolver(const PrimitiveEquationsSolver &solver, int subcycles = 30)
        : solver_(solver), subcycles_(subcycles)
    {
    }
    void step(Real dt_3d, Vec &full_state) const
    {
        check(drf_mode_split_step(solver_.ctx(), dt_3d, subcycles_, full_state.data()));
    }

  private:
    const PrimitiveEquationsSolver &solver_;
    int subcycles_;
};

// MortarSpace (include/dg/mortar.hpp:32-136) for one connection type
// (DRF_CONN_*).  device < 0 keeps the operators on the host only.
class MortarSpace {
  public:
    MortarSpace(int order, int conn_type, int device = 0)
    {
        drf_mortar *m = nullptr;
        check(drf_mortar_create(order, conn_type, device, &m));
        m_.reset(m);
    }
    int num_dofs() const { return drf_mortar_num_dofs(m_.get()); }
    int num_fine_elements() const { return drf_mortar_num_fine(m_.get()); }
    // row-major num_dofs x num_dofs matrices; fine operators stacked
    Vec mass_matrix() const { return op(0); }
    Vec mass_matrix_inv() const { return op(1); }
    Vec projection_coarse() const { return op(2); }
    Vec lift_coarse() const { return op(3); }
    Vec projection_fine() const { return op(4); }
    Vec lift_fine() const { return op(5); }
    // compute_mortar_flux (mortar.hpp:103) for n interfaces, device buffers
    void compute_mortar_flux(int flux_kind, const double *params, int64_t n, const void *Uc,
                             const void *Uf, const void *normals, void *rhs_c, void *rhs_f,
                             void *stream = nullptr)
    {
        check(drf_mortar_flux_batch(m_.get(), flux_kind, params, n, Uc, Uf, normals, rhs_c,
                                    rhs_f, stream));
    }

  private:
    Vec op(int which) const
    {
        const size_t mm = (size_t)num_dofs() * num_dofs();
        Vec v((which >= 4 ? num_fine_elements() : 1) * mm);
        double *p[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        p[which] = v.data();
        check(drf_mortar_operators(m_.get(), p[0], p[1], p[2], p[3], p[4], p[5]));
        return v;
    }
    struct Del {
        void operator()(drf_mortar *m) const { drf_mortar_destroy(m); }
    };
    std::unique_ptr<drf_mortar, Del> m_;
};

// ConformingInterface::compute_flux (include/dg/mortar.hpp:202) for n faces
struct ConformingInterface {
    int order;
    void compute_flux(int flux_kind, const double *params, int64_t n, const void *UL,
                      const void *UR, const void *normals, void *rhs_l, void *rhs_r,
                      void *stream = nullptr) const
    {
        check(drf_conforming_flux_batch(order, flux_kind, params, n, UL, UR, normals, rhs_l,
                                        rhs_r, stream));
    }
};

// GradientErrorEstimator(basis, params) + estimate + the marking of
// AdaptiveMeshRefinement::mark_elements, on device buffers.
// AdaptiveMeshRefinement::transfer_solution on the GPU.  Appends to
// new_solution as the reference does (adapt() hands in the solution vector
// itself, refinement.cpp:471): the elements that are not parents in order,
// then the 8 children of an XYZ projection of each parent in map order.  The
// parents must be ascending, as adapt() builds the map (:452-466).
inline void transfer_solution(int order, const std::vector<Vec> &old_solution,
                              const std::vector<std::pair<int64_t, std::vector<int64_t>>> &refinement_map,
                              std::vector<Vec> &new_solution, int device = 0)
{
    const size_t nd = (size_t)(order + 1) * (order + 1) * (order + 1);
    const int64_t E = (int64_t)old_solution.size();
    Vec U(E * nd);
    for (int64_t e = 0; e < E; ++e) {
        if (old_solution[e].size() != nd)
            throw Error(DRF_ERR_ARG, "transfer_solution: element vector of the wrong size");
        std::copy(old_solution[e].begin(), old_solution[e].end(), U.begin() + e * nd);
    }
    std::vector<int64_t> parents;
    parents.reserve(refinement_map.size());
    for (const auto &entry : refinement_map)
        parents.push_back(entry.first);
    const int64_t R = (int64_t)parents.size(), n_out = E + 7 * R;
    Vec out(n_out * nd);
    check(drf_transfer_solution(order, E, U.data(), R, parents.data(), out.data(), device));
    new_solution.reserve(new_solution.size() + n_out);
    for (int64_t k = 0; k < n_out; ++k)
        new_solution.emplace_back(out.begin() + k * nd, out.begin() + (k + 1) * nd);
}

class GradientErrorEstimator {
  public:
    GradientErrorEstimator(int order, Real refine_threshold = 0.1, Real coarsen_threshold = 0.01,
                           Real smoothness_factor = 0.5, int max_level = 8, int min_level = 0,
                           int device = 0)
        : order_(order)
    {
        drf_estimator *e = nullptr;
        check(drf_estimator_create(order, refine_threshold, coarsen_threshold, smoothness_factor,
                                   max_level, min_level, device, &e));
        e_.reset(e);
    }
    void estimate(int64_t n_elem, const void *field, const void *h_char, const void *levels,
                  void *value, void *mask, void *refine, void *coarsen, void *action,
                  void *mask_out, void *stream = nullptr)
    {
        const int64_t nd = (int64_t)(order_ + 1) * (order_ + 1) * (order_ + 1);
        check(drf_estimate_dev(e_.get(), n_elem, field, nd, h_char, levels, value, mask, refine,
                               coarsen, action, mask_out, nullptr, stream));
    }

  private:
    int order_;
    struct Del {
        void operator()(drf_estimator *e) const { drf_estimator_destroy(e); }
    };
    std::unique_ptr<drf_estimator, Del> e_;
};

} // namespace drifter_gpu
