// sph_amd.hpp -- C++ host classes over the C-ABI of sph_capi.h, shaped like
// the two reference interfaces they replace:
//
//   sph_amd::NeighborhoodSearch  ~ CompactNSearch::NeighborhoodSearch as the
//     reference calls it (SPlisHSPlasH/NeighborhoodSearch.h:8-14), in its
//     GPU_NEIGHBORHOOD_SEARCH form where PointSet::neighbor_list returns a raw
//     `const unsigned*` (Simulation.h:466-473).
//   sph_amd::TimeStepDFSPH       ~ SPH::TimeStepDFSPH (TimeStepDFSPH.h:24-76)
//     together with the fluid/boundary state it steps, device-resident.
//
// Header-only; link with libsph_amd.so. Errors raise sph_amd::Error (the C
// layer itself never throws).
#pragma once
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sph_capi.h"

namespace sph_amd {

struct Error : std::runtime_error {
    int code;
    Error(int c, const char *where) : std::runtime_error(std::string(where) + ": " + sph_last_error()), code(c) {}
};

inline void check(int rc, const char *where) {
    if (rc != SPH_OK) throw Error(rc, where);
}

class NeighborhoodSearch;

// View of one point set (CompactNSearch::PointSet).
class PointSet {
public:
    PointSet(sph_ns_t ns, unsigned index) : ns_(ns), index_(index) {}
    std::size_t n_points() const {
        std::size_t n = 0;
        check(sph_ns_n_points(ns_, index_, &n), "n_points");
        return n;
    }
    unsigned n_neighbors(unsigned point_set, unsigned i) const {
        unsigned c = 0;
        check(sph_ns_n_neighbors(ns_, index_, point_set, i, &c), "n_neighbors");
        return c;
    }
    unsigned neighbor(unsigned point_set, unsigned i, unsigned k) const {
        unsigned j = 0;
        check(sph_ns_neighbor(ns_, index_, point_set, i, k, &j), "neighbor");
        return j;
    }
    const unsigned *neighbor_list(unsigned point_set, unsigned i) const {
        const unsigned *p = nullptr;
        unsigned c = 0;
        check(sph_ns_neighbor_list(ns_, index_, point_set, i, &p, &c), "neighbor_list");
        return p;
    }
    // Permutes lst[0..n_points) in place with the last z_sort().
    template <class T> void sort_field(T *lst) const {
        check(sph_ns_sort_field(ns_, index_, lst, sizeof(T)), "sort_field");
    }
    void *get_user_data() const {
        void *u = nullptr;
        check(sph_ns_user_data(ns_, index_, &u), "get_user_data");
        return u;
    }

private:
    sph_ns_t ns_;
    unsigned index_;
};

class NeighborhoodSearch {
public:
    explicit NeighborhoodSearch(double r, bool erase_empty_cells = false) {
        check(sph_ns_create(r, erase_empty_cells ? 1 : 0, &h_), "NeighborhoodSearch");
    }
    ~NeighborhoodSearch() { sph_ns_destroy(h_); }
    NeighborhoodSearch(const NeighborhoodSearch &) = delete;
    NeighborhoodSearch &operator=(const NeighborhoodSearch &) = delete;

    void set_radius(double r) { check(sph_ns_set_radius(h_, r), "set_radius"); }
    double radius() const {
        double r = 0;
        check(sph_ns_get_radius(h_, &r), "radius");
        return r;
    }
    void reset() {
        check(sph_ns_reset(h_), "reset");
        sets_.clear();
    }
    // x: n points of 3 contiguous Real (float or double); the pointer is kept.
    template <class Real>
    unsigned add_point_set(Real const *x, std::size_t n, bool is_dynamic = true, bool search_neighbors = true,
                           bool find_neighbors = true, void *user_data = nullptr) {
        static_assert(std::is_same<Real, float>::value || std::is_same<Real, double>::value, "float or double");
        unsigned idx = 0;
        check(sph_ns_add_point_set(h_, x, n, std::is_same<Real, double>::value ? SPH_F64 : SPH_F32,
                                   is_dynamic, search_neighbors, find_neighbors, user_data, &idx),
              "add_point_set");
        sets_.emplace_back(h_, idx);
        return idx;
    }
    template <class Real> void resize_point_set(unsigned i, Real const *x, std::size_t n) {
        check(sph_ns_resize_point_set(h_, i, x, n), "resize_point_set");
    }
    void update_point_sets() { check(sph_ns_update_point_sets(h_), "update_point_sets"); }
    void find_neighbors(bool points_changed = true) {
        check(sph_ns_find_neighbors(h_, points_changed ? 1 : 0), "find_neighbors");
    }
    // Point query: neighbours of x in every point set.
    template <class Real> void find_neighbors(Real const *x, std::vector<std::vector<unsigned>> &neighbors) {
        const double q[3] = {(double)x[0], (double)x[1], (double)x[2]};
        neighbors.assign(sets_.size(), {});
        for (unsigned s = 0; s < sets_.size(); s++) {
            unsigned c = 0;
            check(sph_ns_query_point(h_, q, s, nullptr, 0, &c), "find_neighbors(x)");
            neighbors[s].resize(c);
            if (c) check(sph_ns_query_point(h_, q, s, neighbors[s].data(), c, &c), "find_neighbors(x)");
        }
    }
    void z_sort() { check(sph_ns_z_sort(h_), "z_sort"); }
    void set_active(bool active) { check(sph_ns_set_active_all(h_, active), "set_active"); }
    void set_active(unsigned i, bool search_neighbors, bool find_neighbors) {
        check(sph_ns_set_active_set(h_, i, search_neighbors, find_neighbors), "set_active");
    }
    void set_active(unsigned i, unsigned j, bool active) {
        check(sph_ns_set_active_pair(h_, i, j, active), "set_active");
    }
    bool is_active(unsigned i, unsigned j) const {
        int a = 0;
        check(sph_ns_is_active_pair(h_, i, j, &a), "is_active");
        return a != 0;
    }
    std::size_t n_point_sets() const { return sets_.size(); }
    PointSet const &point_set(unsigned i) const { return sets_.at(i); }
    std::vector<PointSet> const &point_sets() const { return sets_; }
    sph_ns_t handle() const { return h_; }

private:
    sph_ns_t h_ = nullptr;
    std::vector<PointSet> sets_;
};

// Device-resident DFSPH time step. The member names follow
// SPH::TimeStepDFSPH; each sub-step is also callable on its own.
class TimeStepDFSPH {
public:
    explicit TimeStepDFSPH(const sph_dfsph_params_t &p) { check(sph_dfsph_create(&p, &h_), "TimeStepDFSPH"); }
    static sph_dfsph_params_t default_params() {
        sph_dfsph_params_t p;
        sph_dfsph_default_params(&p);
        return p;
    }
    ~TimeStepDFSPH() { sph_dfsph_destroy(h_); }
    TimeStepDFSPH(const TimeStepDFSPH &) = delete;
    TimeStepDFSPH &operator=(const TimeStepDFSPH &) = delete;

    void set_fluid(const float *x, const float *v, std::size_t n) {
        check(sph_dfsph_set_fluid(h_, x, v, n), "set_fluid");
    }
    void add_boundary_particles(const float *x, std::size_t n) {
        check(sph_dfsph_add_boundary_particles(h_, x, n), "add_boundary_particles");
    }
    void set_volume_map_box(const double lo[3], const double hi[3], const unsigned res[3], bool invert = true,
                            double thickness = 0.0) {
        check(sph_dfsph_set_volume_map_box(h_, lo, hi, invert, thickness, res), "set_volume_map_box");
    }
    void add_volume_map_sphere(const double centre[3], double radius, const unsigned res[3], bool invert = false,
                               double thickness = 0.0) {
        check(sph_dfsph_add_volume_map_sphere(h_, centre, radius, invert, thickness, res), "add_volume_map_sphere");
    }
    // Learned warm start (no reference counterpart): solver 0 pressure, 1
    // divergence; theta = 4673 floats, or nullptr for the reference's start.
    void set_warm_start_model(int solver, const float *theta) {
        check(sph_dfsph_set_warm_start_model(h_, solver, theta), "set_warm_start_model");
    }
    void set_warm_start_mode(int solver, int mode) {
        check(sph_dfsph_set_warm_start_mode(h_, solver, mode), "set_warm_start_mode");
    }
    void set_warm_start_policy(int solver, int min_prev_iterations) {
        check(sph_dfsph_set_warm_start_policy(h_, solver, min_prev_iterations), "set_warm_start_policy");
    }
    void init() { check(sph_dfsph_init(h_), "init"); }
    void reset() { check(sph_dfsph_reset(h_), "reset"); }
    // Emitter support: emittedParticles + resize of the reference's TimeStep.
    void emittedParticles(const float *x, const float *v, std::size_t n) {
        check(sph_dfsph_emit_particles(h_, x, v, n), "emit_particles");
    }
    // EmitterSystem on the device (EmitterSystem.cpp:93-117): addEmitter with
    // the emitter's start and end times, enableReuseParticles /
    // disableReuseParticles, and the fluid's maxEmitterParticles.
    void addEmitter(unsigned width, unsigned height, const double pos[3], const double rotation[9], double velocity,
                    int type, double emit_start_time = 0.0, double emit_end_time = 1.7976931348623157e308) {
        sph_emitter_t e{};
        e.type = type;
        e.width = width;
        e.height = height;
        for (int c = 0; c < 3; c++) e.translation[c] = pos[c];
        for (int c = 0; c < 9; c++) e.rotation[c] = rotation[c];
        e.velocity = velocity;
        e.emit_start_time = emit_start_time;
        e.emit_end_time = emit_end_time;
        check(sph_dfsph_add_emitter(h_, &e), "addEmitter");
    }
    void enableReuseParticles(const double box_min[3], const double box_max[3]) {
        check(sph_dfsph_set_emitter_reuse(h_, 1, box_min, box_max), "enableReuseParticles");
    }
    void disableReuseParticles() { check(sph_dfsph_set_emitter_reuse(h_, 0, nullptr, nullptr), "disableReuseParticles"); }
    void setMaxEmitterParticles(std::size_t n) {
        check(sph_dfsph_set_max_emitter_particles(h_, n), "setMaxEmitterParticles");
    }
    void step() { check(sph_dfsph_step(h_, 1), "step"); }
    void step(int n) { check(sph_dfsph_step(h_, n), "step"); }
    // sub-steps, TimeStepDFSPH.cpp:117-249
    void performNeighborhoodSearch() { stage(SPH_STAGE_NEIGHBORHOOD_SEARCH); }
    void computeVolumeAndBoundaryX() { stage(SPH_STAGE_VOLUME_MAPS); }
    void computeDensities() { stage(SPH_STAGE_DENSITIES); }
    void computeDFSPHFactor() { stage(SPH_STAGE_DFSPH_FACTOR); }
    void divergenceSolve() { stage(SPH_STAGE_DIVERGENCE_SOLVE); }
    void clearAccelerations() { stage(SPH_STAGE_CLEAR_ACCEL); }
    void computeViscosity() { stage(SPH_STAGE_VISCOSITY); }
    void computeXSPH() { stage(SPH_STAGE_XSPH); }
    void updateTimeStepSize() { stage(SPH_STAGE_CFL); }
    void pressureSolve() { stage(SPH_STAGE_PRESSURE_SOLVE); }
    void stage(int s) { check(sph_dfsph_stage(h_, s), "stage"); }

    std::size_t numParticles() const {
        std::size_t n = 0;
        check(sph_dfsph_num_particles(h_, &n), "num_particles");
        return n;
    }
    // Field access by the reference field names, float32, particle-id order.
    std::vector<float> field(const char *name, int components) const {
        std::vector<float> out(numParticles() * components);
        check(sph_dfsph_get_field(h_, name, out.data(), out.size() * sizeof(float)), name);
        return out;
    }
    void set_field(const char *name, const std::vector<float> &v) {
        check(sph_dfsph_set_field(h_, name, v.data(), v.size() * sizeof(float)), name);
    }
    sph_dfsph_counters_t counters() const {
        sph_dfsph_counters_t c;
        check(sph_dfsph_get_counters(h_, &c), "counters");
        return c;
    }
    unsigned getNumIterations() const { return (unsigned)counters().iterations; }
    sph_dfsph_t handle() const { return h_; }

private:
    sph_dfsph_t h_ = nullptr;
};

}  // namespace sph_amd
