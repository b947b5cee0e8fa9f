// C++ host classes with the reference's solver surface, over the C ABI of
// include/netflux.h.  Header-only; link with libnetflux.so.
//
//   netflux::Config        the option and run-state subset of Config
//                          (include/Config.hpp:8-175) the step depends on
//   netflux::EulerSolver   drop-in for EulerSolver (include/Solver.hpp:52-79):
//                          step(const Config&), calc_timestep(Config&),
//                          get_solver_data(), get_solver_type()
//   netflux::SolverData    get_solution(), get_primvars(),
//                          get_primvars_gradient(), get_primvars_limiter(),
//                          get_flux_balance() as AoS VecField copies
//   netflux::Driver        the time loop of src/Driver.cpp:31-57 (no VTK)
//   netflux::Mesh          owner of a library mesh: SU2 files
//                          (PrimalGrid::read_su2_mesh, src/geometry/
//                          PrimalGrid.cpp:140-244) or generated boxes
//
// Errors are thrown as std::runtime_error carrying the reference's messages
// (src/Solver.cpp:267-268, src/SolverData.cpp:73-78).
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/netflux.h"

namespace netflux {

enum class TimeScheme { ExplicitEuler, TVD_RK3 };
enum class SpatialOrder { First, Second };
enum class Limiter { NONE, Barth };
enum class InviscidFluxScheme { Rusanov, HLLC };
enum class BoundaryType { NoSlipWall, SlipWall, FarField };

struct Config {
    TimeScheme time_scheme = TimeScheme::TVD_RK3;
    SpatialOrder spatial_order = SpatialOrder::Second;
    Limiter limiter = Limiter::Barth;
    InviscidFluxScheme inv_flux_scheme = InviscidFluxScheme::Rusanov;
    double CFL = 0.5;
    double primvars_inf[5] = {1.225, 0.0, 0.0, 0.0, 101325.0};  // standard air defaults
    bool check_physical_validity = false;
    size_t n_timesteps = 1;
    bool rk3_zero_old_first_step = true;
    int precision = NF_PREC_FP64;
    int cell_order = NF_ORDER_MORTON;
    bool local_time_step = false;  // per-cell CFL step for steady runs (not the reference)
    // run state (Config::delta_time / time / timestep)
    double delta_time = 0.0;
    double time = 0.0;
    size_t timestep = 0;

    double get_delta_time() const { return delta_time; }
    void set_delta_time(double dt) { delta_time = dt; }
    double get_time() const { return time; }
    void set_time(double t) { time = t; }
    size_t get_timestep() const { return timestep; }
    void set_timestep(size_t n) { timestep = n; }
    size_t get_n_timesteps() const { return n_timesteps; }

    nf_params to_params() const {
        nf_params p{};
        p.time_scheme = time_scheme == TimeScheme::TVD_RK3 ? NF_TVD_RK3 : NF_EXPLICIT_EULER;
        p.spatial_order = spatial_order == SpatialOrder::Second ? NF_SECOND_ORDER : NF_FIRST_ORDER;
        p.limiter = limiter == Limiter::Barth ? NF_LIMITER_BARTH : NF_LIMITER_NONE;
        p.inv_flux_scheme = inv_flux_scheme == InviscidFluxScheme::Rusanov ? NF_FLUX_RUSANOV : NF_FLUX_HLLC;
        p.CFL = CFL;
        for (int k = 0; k < 5; ++k) p.primvars_inf[k] = primvars_inf[k];
        p.check_physical_validity = check_physical_validity;
        p.rk3_zero_old_first_step = rk3_zero_old_first_step;
        p.precision = precision;
        p.cell_order = cell_order;
        p.local_time_step = local_time_step;
        return p;
    }
};

// AoS [cells][rows] doubles, the layout of VecField / GradField
struct VecField {
    uint32_t n = 0, rows = 0;
    std::vector<double> data;
    double operator()(uint32_t i, uint32_t k) const { return data[(size_t)i * rows + k]; }
    uint32_t size() const { return n; }
};

inline void check(int rc, const nf_ctx *ctx, const char *what) {
    if (rc == NF_OK) return;
    const std::string msg = ctx ? nf_last_error(ctx) : "";
    if (rc == NF_E_INVALID_DT || rc == NF_E_NONPHYSICAL) throw std::runtime_error(msg);
    throw std::runtime_error(std::string(what) + " failed (status " + std::to_string(rc) + "): " + msg);
}

class EulerSolver;

class SolverData {
    const EulerSolver *s_;

  public:
    explicit SolverData(const EulerSolver *s) : s_(s) {}
    VecField get_solution() const;
    VecField get_primvars() const;
    VecField get_primvars_gradient() const;  // column-major 5x3 per cell
    VecField get_primvars_limiter() const;
    VecField get_flux_balance() const;
};

class EulerSolver {
    nf_ctx *ctx_ = nullptr;
    nf_fv_sizes sizes_{};
    SolverData data_{this};

  public:
    EulerSolver(const Config &config, const nf_mesh_desc &mesh) {
        const nf_params p = config.to_params();
        check(nf_create(&mesh, &p, &ctx_), nullptr, "nf_create");
        check(nf_ctx_sizes(ctx_, &sizes_), ctx_, "nf_ctx_sizes");
    }
    ~EulerSolver() { nf_destroy(ctx_); }
    EulerSolver(const EulerSolver &) = delete;
    EulerSolver &operator=(const EulerSolver &) = delete;

    void calc_timestep(Config &config) {
        double dt = 0.0;
        check(nf_calc_timestep(ctx_, &dt), ctx_, "nf_calc_timestep");
        config.set_delta_time(dt);
    }
    void step(const Config &config) { check(nf_step(ctx_, config.get_delta_time()), ctx_, "nf_step"); }
    const SolverData &get_solver_data() const { return data_; }
    const char *get_solver_type() const { return "Euler"; }

    // additions: state injection and the device-resident loop
    void set_state(const std::vector<double> &U, const std::vector<double> *U_old = nullptr) {
        check(nf_set_state(ctx_, U.data(), U_old ? U_old->data() : nullptr), ctx_, "nf_set_state");
    }
    nf_run_stats run(int n_steps, double residual_drop = 0.0, int check_every = 10) {
        nf_run_opts o{n_steps, residual_drop, check_every, 1};
        nf_run_stats st{};
        check(nf_run(ctx_, &o, &st), ctx_, "nf_run");
        return st;
    }
    void eval_flux_balance() { check(nf_eval_flux_balance(ctx_), ctx_, "nf_eval_flux_balance"); }
    uint32_t n_cells() const { return sizes_.n_interior_cells; }
    nf_ctx *handle() const { return ctx_; }
};

inline VecField fetch(const EulerSolver *s, int which, uint32_t rows) {
    VecField f;
    f.n = s->n_cells();
    f.rows = rows;
    f.data.resize((size_t)f.n * rows);
    if (which < 0) check(nf_get_state(s->handle(), f.data.data()), s->handle(), "nf_get_state");
    else check(nf_get_field(s->handle(), which, f.data.data()), s->handle(), "nf_get_field");
    return f;
}
inline VecField SolverData::get_solution() const { return fetch(s_, -1, 5); }
inline VecField SolverData::get_primvars() const { return fetch(s_, NF_FIELD_PRIM, 5); }
inline VecField SolverData::get_primvars_gradient() const { return fetch(s_, NF_FIELD_GRADIENT, 15); }
inline VecField SolverData::get_primvars_limiter() const { return fetch(s_, NF_FIELD_LIMITER, 5); }
inline VecField SolverData::get_flux_balance() const { return fetch(s_, NF_FIELD_RESIDUAL, 5); }

// Several meshes advanced together, one launch per sweep over all of them
// (no reference counterpart: the reference asserts a single solver,
// src/Driver.cpp:36).  Only the freestream may differ between the configs.
class BatchSolver {
    nf_ctx *ctx_ = nullptr;
    std::vector<uint32_t> n_;

  public:
    BatchSolver(const std::vector<Config> &configs, const std::vector<nf_mesh_desc> &meshes) {
        if (configs.size() != meshes.size() || meshes.empty()) throw std::invalid_argument("one config per mesh");
        std::vector<nf_params> p;
        for (const Config &c : configs) p.push_back(c.to_params());
        check(nf_create_batch((int32_t)meshes.size(), meshes.data(), p.data(), &ctx_), nullptr, "nf_create_batch");
        for (size_t i = 0; i < meshes.size(); ++i) {
            nf_fv_sizes s{};
            check(nf_inst_sizes(ctx_, (int32_t)i, &s), ctx_, "nf_inst_sizes");
            n_.push_back(s.n_interior_cells);
        }
    }
    ~BatchSolver() { nf_destroy(ctx_); }
    BatchSolver(const BatchSolver &) = delete;
    BatchSolver &operator=(const BatchSolver &) = delete;

    size_t size() const { return n_.size(); }
    uint32_t n_cells(size_t i) const { return n_[i]; }
    void set_state(size_t i, const std::vector<double> &U, const std::vector<double> *U_old = nullptr) {
        check(nf_inst_set_state(ctx_, (int32_t)i, U.data(), U_old ? U_old->data() : nullptr), ctx_,
              "nf_inst_set_state");
    }
    std::vector<double> get_solution(size_t i) const {
        std::vector<double> U((size_t)5 * n_[i]);
        check(nf_inst_get_state(ctx_, (int32_t)i, U.data()), ctx_, "nf_inst_get_state");
        return U;
    }
    std::vector<nf_run_stats> run(int n_steps, double residual_drop = 0.0, int check_every = 10) {
        nf_run_opts o{n_steps, residual_drop, check_every, 1, 0, 0.0};
        std::vector<nf_run_stats> st(n_.size());
        check(nf_run_batch(ctx_, &o, st.data()), ctx_, "nf_run_batch");
        return st;
    }
    nf_tune_result autotune(int steps = 10) {
        nf_tune_result r{};
        check(nf_autotune(ctx_, steps, &r), ctx_, "nf_autotune");
        return r;
    }
    nf_ctx *handle() const { return ctx_; }
};

// A library-owned primal mesh.  read_su2 maps each marker tag to a boundary
// type by name; an unnamed marker or a non-tetrahedral element throws with
// the reader's message.
class Mesh {
    nf_mesh *m_ = nullptr;

  public:
    explicit Mesh(nf_mesh *m) : m_(m) {}
    Mesh(const Mesh &) = delete;
    Mesh &operator=(const Mesh &) = delete;
    Mesh(Mesh &&o) noexcept : m_(o.m_) { o.m_ = nullptr; }
    ~Mesh() {
        if (m_) nf_mesh_destroy(m_);
    }
    static Mesh read_su2(const std::string &path, const std::vector<std::string> &names,
                         const std::vector<BoundaryType> &types) {
        if (names.size() != types.size()) throw std::runtime_error("read_su2: one boundary type per patch name");
        std::vector<const char *> n;
        std::vector<int32_t> bc;
        for (size_t i = 0; i < names.size(); ++i) {
            n.push_back(names[i].c_str());
            bc.push_back(types[i] == BoundaryType::NoSlipWall ? NF_BC_NOSLIP_WALL
                         : types[i] == BoundaryType::SlipWall ? NF_BC_SLIP_WALL
                                                              : NF_BC_FARFIELD);
        }
        nf_mesh *m = nullptr;
        if (nf_mesh_read_su2(path.c_str(), (int32_t)n.size(), n.data(), bc.data(), &m) != NF_OK)
            throw std::runtime_error("Error reading su2 mesh '" + path + "', " + nf_mesh_last_error());
        return Mesh(m);
    }
    nf_mesh_desc desc() const {
        nf_mesh_desc d;
        if (nf_mesh_get_desc(m_, &d) != NF_OK) throw std::runtime_error("nf_mesh_get_desc failed");
        return d;
    }
    nf_mesh *handle() const { return m_; }
};

// src/Driver.cpp:31-57 without output
class Driver {
    Config &config_;
    EulerSolver solver_;

  public:
    Driver(Config &config, const nf_mesh_desc &mesh) : config_(config), solver_(config, mesh) {}
    EulerSolver &solver() { return solver_; }
    void solve() {
        config_.set_timestep(0);
        config_.set_time(0.0);
        while (true) {
            solver_.calc_timestep(config_);
            solver_.step(config_);
            config_.set_timestep(config_.get_timestep() + 1);
            config_.set_time(config_.get_time() + config_.get_delta_time());
            if (config_.get_timestep() >= config_.get_n_timesteps()) break;
        }
    }
};

}  // namespace netflux
