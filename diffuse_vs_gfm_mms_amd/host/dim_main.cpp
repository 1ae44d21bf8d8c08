// dim_main.cpp — DIM case driver on the GPU library.
//
//     dim_main_<D>d <config_file> [--precision fp64|fp32_mixed] [--graph] [--device-ic]
//
// The reference's mm_main (src/app/multimaterial_main.cpp:17-62) for
// `interface_method = DIM` case files: the same command line, case-file format,
// console lines, solution CSVs, runtime report and conservation report
// (SOLVER_CONSERVATION, SOLVER_CONSERVATION_INTERVAL), and QUANT_TIMING_ONLY.
// The time loop of run_dim_case (src/app/solver/run_dim.hpp:18-186) runs on
// the device through include/dimhip.h: the state stays in HBM, the loop
// advances to each output time with dt and time on the device, and the host
// receives the CSV row values only at output times.
//
// The precision form is chosen outside the case file (the format stays the
// reference's): --precision, or DIMHIP_PRECISION in the environment; the
// default is the FP64 form, which reproduces the reference bit for bit.
// --device-ic generates the initial state on the GPU (dimhip_initialise)
// instead of building primitive rows on the host; its tanh is the device
// library's, so diffuse initial cells may differ from the reference's in the
// last bit.
//
// Two environment variables reach what the case format cannot express:
//   DIM_MAX_STEPS=<n>         stop after n steps of each resolution (before
//                             tfinal if it comes first), then write the
//                             end-of-run outputs as at tfinal; the reference
//                             has no step-cap key (run_dim.hpp:88).
//   DIMHIP_DUMP_INITIAL=<f>   write the packed FP64 initial state (ncells x
//                             (2 nmat + dim) doubles, the reference's per-cell
//                             order) of each resolution to <f>.<index>.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "config.hpp"
#include "dimhip.h"
#include "initial_conditions.hpp"
#include "output.hpp"

#ifndef DIM_
#error "build with -DDIM_=1, 2 or 3"
#endif

namespace dimhost {

struct Options {
    std::string config;
    int precision = DIMHIP_FP64;
    bool graph = false;
    bool device_ic = false;
};

inline void check(int status, const dimhip_ctx* ctx)
{
    if (status != 0) throw std::runtime_error(dimhip_last_error(ctx));
}

// DIM_MAX_STEPS: a non-negative step count, or -1 when unset
inline long long env_max_steps()
{
    const char* v = std::getenv("DIM_MAX_STEPS");
    if (!v) return -1;
    char* end = nullptr;
    const long long n = std::strtoll(v, &end, 10);
    if (end == v || *end != '\0' || n < 0) throw std::runtime_error("DIM_MAX_STEPS must be a non-negative integer");
    return n;
}

inline bool env_flag(const char* name)
{
    const char* v = std::getenv(name);
    if (!v) return false;
    const std::string f(v);
    return f == "1" || f == "true" || f == "TRUE" || f == "on" || f == "ON" || f == "yes" || f == "YES";
}

inline int env_interval()
{
    const char* v = std::getenv("SOLVER_CONSERVATION_INTERVAL");
    if (!v) return 1;
    try {
        return std::max(1, std::stoi(v));
    } catch (...) {
        return 1;
    }
}

inline int bc_code(const std::string& s)
{
    if (s == "reflective") return DIMHIP_BC_REFLECTIVE;
    if (s == "reflection") return DIMHIP_BC_REFLECTION;
    if (s == "nonreflective" || s == "non_reflective") return DIMHIP_BC_NONREFLECTIVE;
    return DIMHIP_BC_TRANSMISSIVE;
}

// build_dim_material_params (src/app/material/material_builder.hpp:11-48,149-167)
// into the library descriptor.
template<int DIM>
dimhip_desc make_desc(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, int precision)
{
    dimhip_desc d{};
    d.dim = DIM;
    for (int a = 0; a < 3; ++a) {
        d.N[a] = 1;
        d.dx[a] = 1.0;
    }
    for (int a = 0; a < DIM; ++a) {
        d.N[a] = N[a];
        d.dx[a] = (c.domain_max[a] - c.domain_min[a]) / static_cast<double>(N[a]);
        d.bc_lo[a] = bc_code(c.bc_lo[a]);
        d.bc_hi[a] = bc_code(c.bc_hi[a]);
    }
    const int nmat = (int)c.materials.size();
    if (nmat > DIMHIP_MAX_MAT) throw std::runtime_error("dimhip: at most 4 materials");
    d.nmat = nmat;
    for (const auto& m : c.materials) {
        if (m.id < 0 || m.id >= nmat) throw std::runtime_error("build_dim_material_params: invalid material id");
        const bool noble_abel = m.type == "real_gas" || m.type == "noble_abel" || m.type == "noble_abel_gas";
        if (m.type == "ideal_gas") d.eos_kind[m.id] = DIMHIP_EOS_IDEAL;
        else if (m.type == "stiffened_gas") d.eos_kind[m.id] = DIMHIP_EOS_STIFFENED;
        else if (noble_abel) d.eos_kind[m.id] = DIMHIP_EOS_NOBLE_ABEL;
        else if (m.type == "peng_robinson" || m.type == "peng-robinson" || m.type == "tait" ||
                 m.type == "mie_gruneisen" || m.type == "mie-gruneisen" || m.type == "mie_gruneisen_fluid")
            throw std::runtime_error("dimhip: EOS type " + m.type +
                                     " is not supported on the GPU path (ideal_gas, stiffened_gas, noble_abel)");
        else throw std::runtime_error("Unsupported EOS type: " + m.type);
        const auto g = m.params.find("gamma");
        if (g == m.params.end()) throw std::runtime_error("build_material_params: missing gamma in material");
        d.gamma[m.id] = g->second;
        d.p_inf[m.id] = 0.0;
        if (m.type == "stiffened_gas") {
            const auto p = m.params.find("p_inf");
            if (p == m.params.end())
                throw std::runtime_error("build_material_params: missing p_inf in stiffened gas material");
            d.p_inf[m.id] = p->second;
        }
        // co-volume: key b, else covolume, else 0 (material_builder.hpp:31-38)
        d.covolume[m.id] = 0.0;
        if (noble_abel) {
            if (m.params.count("b")) d.covolume[m.id] = m.params.at("b");
            else if (m.params.count("covolume")) d.covolume[m.id] = m.params.at("covolume");
        }
    }
    d.lambda_model = c.dim_lambda_model == "allaire"          ? DIMHIP_LAMBDA_ALLAIRE
                     : c.dim_lambda_model == "equal_velocity" ? DIMHIP_LAMBDA_EQUAL_VELOCITY
                                                              : DIMHIP_LAMBDA_KAPILA;
    d.alpha_source_floor = c.dim_alpha_source_floor;
    d.time_update = c.time_update == "unsplit" ? DIMHIP_UNSPLIT : DIMHIP_SPLIT;
    d.precision = precision;
    d.device = 0;
    return d;
}

// The case's initial condition for dimhip_initialise (material ids -> slots,
// material_slot_from_id, dim_initial_conditions.hpp:170-183).
template<int DIM>
dimhip_ic make_ic(const CaseConfig<DIM>& c)
{
    dimhip_ic ic{};
    const std::string& k = c.initial_condition;
    if (k == "regions") ic.kind = DIMHIP_IC_REGIONS;
    else if (k == "planar_regions") ic.kind = DIMHIP_IC_PLANAR_REGIONS;
    else if (k == "shock_bubble") ic.kind = DIMHIP_IC_SHOCK_BUBBLE;
    else if (k == "coated_shock_bubble") ic.kind = DIMHIP_IC_COATED_SHOCK_BUBBLE;
    else throw std::runtime_error("dim::initialise_dim_from_config: unsupported initial_condition: " + k);
    const auto slot = [&](int id) {
        for (int s = 0; s < (int)c.materials.size(); ++s)
            if (c.materials[s].id == id) return s;
        throw std::runtime_error("dim::material_slot_from_id: material id not found");
    };
    for (int d = 0; d < DIM; ++d) {
        ic.domain_min[d] = c.domain_min[d];
        ic.domain_max[d] = c.domain_max[d];
        ic.planar_normal[d] = c.planar_normal[d];
        ic.bubble_center[d] = c.bubble_center[d];
    }
    ic.interface_thickness = c.interface_thickness;
    ic.sharpness_alpha = c.interface_sharpness_alpha;
    if (ic.kind == DIMHIP_IC_REGIONS || ic.kind == DIMHIP_IC_PLANAR_REGIONS) {
        if (c.regions.size() > DIMHIP_MAX_REGIONS) throw std::runtime_error("dimhip_initialise: too many regions");
        ic.nregions = (int)c.regions.size();
        for (size_t i = 0; i < c.regions.size(); ++i) {
            const auto& r = c.regions[i];
            auto& o = ic.regions[i];
            for (int d = 0; d < DIM; ++d) {
                o.lower[d] = r.lower[d];
                o.upper[d] = r.upper[d];
                o.vel[d] = r.vel[d];
            }
            o.rho = r.rho;
            o.p = r.p;
            o.material = slot(r.material_id);
        }
    } else {
        const auto state = [&](const PhaseState<DIM>& ps, dimhip_ic_state& o) {
            o.rho = ps.rho;
            o.p = ps.p;
            for (int d = 0; d < DIM; ++d) o.vel[d] = ps.vel[d];
            o.material = slot(ps.material);
        };
        state(c.left, ic.left);
        state(c.right, ic.right);
        state(c.bubble, ic.bubble);
        if (ic.kind == DIMHIP_IC_COATED_SHOCK_BUBBLE) state(c.film, ic.film);
        else ic.film.material = -1;
        ic.shock_axis = c.shock_axis;
        ic.shock_position = c.shock_position;
        ic.bubble_radius = c.bubble_radius;
        ic.film_radius = c.film_radius;
    }
    return ic;
}

class Context {
public:
    explicit Context(const dimhip_desc& d)
    {
        check(dimhip_create(&d, &ctx_), nullptr);
    }
    ~Context() { dimhip_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    dimhip_ctx* get() const { return ctx_; }

private:
    dimhip_ctx* ctx_ = nullptr;
};

// run_dim_case (run_dim.hpp:18-186)
template<int DIM>
void run_case(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, const Options& opt, int n_index)
{
    if (c.barton_solid_material >= 0)
        throw std::runtime_error(
            "run_dim_case: base DIM fluid solver does not run Barton solids; use the Barton DIM extension");
    const dimhip_desc desc = make_desc<DIM>(c, N, opt.precision);
    const int nmat = desc.nmat;
    Context ctx(desc);
    long long ncells = 1;
    for (int d = 0; d < DIM; ++d) ncells *= N[d];
    if (opt.device_ic) {
        const dimhip_ic ic = make_ic<DIM>(c);
        check(dimhip_initialise(ctx.get(), &ic), ctx.get());
    } else {
        const InitialState<DIM> ic(c, N, nmat);
        const std::vector<double> prim = ic.rows();
        check(dimhip_upload_primitive(ctx.get(), prim.data(), (size_t)ncells), ctx.get());
    }
    if (const char* dump = std::getenv("DIMHIP_DUMP_INITIAL")) {
        std::vector<double> packed((size_t)ncells * (2 * nmat + DIM));
        check(dimhip_download(ctx.get(), packed.data(), (size_t)ncells), ctx.get());
        const std::string path = std::string(dump) + "." + std::to_string(n_index);
        FILE* f = std::fopen(path.c_str(), "wb");
        if (!f || std::fwrite(packed.data(), sizeof(double), packed.size(), f) != packed.size())
            throw std::runtime_error("DIMHIP_DUMP_INITIAL: cannot write " + path);
        std::fclose(f);
    }
    const long long max_steps = env_max_steps();
    const bool timing_only = std::getenv("QUANT_TIMING_ONLY") != nullptr;
    const bool track = env_flag("SOLVER_CONSERVATION");
    const int interval = env_interval();
    const int nt = nmat + 1 + DIM + 1;
    const std::string precision = opt.precision == DIMHIP_FP64 ? "fp64" : "fp32_mixed";
    const bool use_graph = opt.graph;
    check(dimhip_set_timing(ctx.get(), use_graph ? 0 : 1), ctx.get());

    double time = 0.0;
    long long step = 0;
    size_t next_output = 0;
    PhaseTimes ph;
    const auto now = [] { return std::chrono::steady_clock::now(); };
    const auto secs = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
    const auto wall_start = now();

    std::vector<double> rows;
    const auto snapshot = [&](const std::string& suffix) {
        const auto t0 = now();
        rows.resize((size_t)ncells * (3 + DIM + 3 * nmat));
        check(dimhip_output_rows(ctx.get(), rows.data(), (size_t)ncells), ctx.get());
        write_solution_csv<DIM>(solution_filename<DIM>(c, N, suffix), c, N, nmat, rows);
        ph.output += secs(t0, now());
    };
    const auto totals = [&] {
        Totals t;
        t.v.assign(nt, 0.0);
        t.nmat = nmat;
        t.dim = DIM;
        check(dimhip_conservation(ctx.get(), t.v.data()), ctx.get());
        return t;
    };
    std::unique_ptr<ConservationLog> log;
    std::vector<double> flux_integral(nt, 0.0);
    if (track) {
        const auto t0 = now();
        check(dimhip_track_boundary_flux(ctx.get(), 1), ctx.get());
        log = std::make_unique<ConservationLog>(report_filename<DIM>(c, N, "_conservation.csv"), totals());
        log->write(step, time, totals(), flux_integral);
        ph.conservation += secs(t0, now());
    }

    while (next_output < c.output_times.size() && c.output_times[next_output] <= 1e-14)
        snapshot(time_tag(c.output_times[next_output++]));

    while (time < c.tfinal - 1e-14 && (max_steps < 0 || step < max_steps)) {
        const double target = next_output < c.output_times.size() ? c.output_times[next_output] : c.tfinal;
        // steps until the next output time (the device ends the run of steps there),
        // until the next conservation sample, or a bounded batch whose kernel
        // timings are then collected
        int chunk = use_graph ? 1 << 20 : 2048;
        if (track) chunk = std::min<long long>(chunk, interval - step % interval);
        if (max_steps >= 0) chunk = (int)std::min<long long>(chunk, max_steps - step);
        long long taken = 0;
        check(dimhip_advance(ctx.get(), chunk, c.cfl, target, use_graph ? 1 : 0, &time, &taken), ctx.get());
        step += taken;
        if (taken == 0) throw std::runtime_error("run_dim_case: non-positive timestep");
        while (next_output < c.output_times.size() && time >= c.output_times[next_output] - 1e-14)
            snapshot(time_tag(c.output_times[next_output++]));
        const bool last = time >= c.tfinal - 1e-14 || (max_steps >= 0 && step >= max_steps);
        if (log && (step % interval == 0 || last)) {
            const auto t0 = now();
            check(dimhip_boundary_flux_integral(ctx.get(), flux_integral.data()), ctx.get());
            log->write(step, time, totals(), flux_integral);
            ph.conservation += secs(t0, now());
        }
    }
    check(dimhip_synchronize(ctx.get()), ctx.get());
    const double wall = secs(wall_start, now());
    if (!use_graph) {
        double ms[3] = {0, 0, 0};
        long long cnt[3] = {0, 0, 0};
        check(dimhip_kernel_times(ctx.get(), ms, cnt), ctx.get());
        ph.sweep = ms[0] * cnt[0] * 1e-3;
        ph.boundary = ms[1] * cnt[1] * 1e-3;
        ph.cfl = ms[2] * cnt[2] * 1e-3;
    }
    if (c.output_times.empty() && !timing_only) snapshot("");
    write_runtime_report<DIM>(c, N, step, time, wall, ph, precision);
}

inline Options parse_args(int argc, char** argv)
{
    if (argc < 2) throw std::runtime_error("Usage: ./multimaterial_main <config_file>");
    Options o;
    o.config = argv[1];
    if (const char* p = std::getenv("DIMHIP_PRECISION")) {
        const std::string v(p);
        if (v == "fp32_mixed") o.precision = DIMHIP_FP32_MIXED;
        else if (v != "fp64") throw std::runtime_error("DIMHIP_PRECISION must be fp64 or fp32_mixed");
    }
    for (int i = 2; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--graph") o.graph = true;
        else if (a == "--device-ic") o.device_ic = true;
        else if (a == "--precision" && i + 1 < argc) {
            const std::string v = argv[++i];
            if (v == "fp64") o.precision = DIMHIP_FP64;
            else if (v == "fp32_mixed") o.precision = DIMHIP_FP32_MIXED;
            else throw std::runtime_error("--precision must be fp64 or fp32_mixed");
        } else throw std::runtime_error("unknown argument: " + a);
    }
    return o;
}

}  // namespace dimhost

int main(int argc, char** argv)
{
    using namespace dimhost;
    try {
        const Options opt = parse_args(argc, argv);
        if (selects_barton(opt.config))
            throw std::runtime_error("the Barton DIM solid extension is not part of this driver");
        const CaseConfig<DIM_> c = load_case<DIM_>(opt.config);
        std::cout << "Regions parsed: " << c.regions.size() << "\n";
        std::cout << "Interface method: " << c.interface_method << "\n";
        for (size_t n_index = 0; n_index < c.N_list.size(); ++n_index) {
            const auto& N = c.N_list[n_index];
            std::cout << "Running N = ";
            for (int d = 0; d < DIM_; ++d) std::cout << N[d] << " ";
            std::cout << "\n";
            if (c.interface_method == "DIM") run_case<DIM_>(c, N, opt, (int)n_index);
            else if (c.interface_method == "GFM" || c.interface_method == "SM")
                throw std::runtime_error("interface_method " + c.interface_method +
                                         " runs on the reference's sharp-interface solver, not this DIM driver");
            else throw std::runtime_error("Unknown interface_method: " + c.interface_method);
        }
        std::cout << "Finished.\n";
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
