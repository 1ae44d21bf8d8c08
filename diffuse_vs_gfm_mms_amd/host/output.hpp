// output.hpp — solution CSVs, the runtime report and the conservation report
// in the reference's formats.
//
// Files, names and number formatting follow src/io/dim_write_csv.hpp:14-124,
// src/app/io/dim_output_utils.hpp:17-95, src/app/io/runtime_report.hpp:14-149
// and src/app/io/conservation_report.hpp:53-119,407-698.  Numbers are written
// as the reference's default-formatted std::ostream writes them (printf
// "%g", 6 significant digits).  The per-cell CSV values come from the device
// (dimhip_output_rows); the host only formats them, in parallel blocks
// written in order.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "config.hpp"

namespace dimhost {

// "%g" with the default 6 significant digits: what `stream << double` prints.
inline void put_number(std::string& out, double v)
{
    char buf[40];
    const int n = std::snprintf(buf, sizeof buf, "%g", v);
    out.append(buf, (size_t)n);
}

inline std::string number(double v)
{
    std::string s;
    put_number(s, v);
    return s;
}

// format_time_tag (dim_output_utils.hpp:25-46): "t" + %.6e with '.'/'+' -> 'p', '-' -> 'm'.
inline std::string time_tag(double t)
{
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6e", t);
    std::string s = buf;
    for (char& ch : s) {
        if (ch == '.' || ch == '+') ch = 'p';
        else if (ch == '-') ch = 'm';
    }
    return "t" + s;
}

template<int DIM>
std::string grid_suffix(const std::array<int, DIM>& N)
{
    std::string s;
    for (int d = 0; d < DIM; ++d) s += "_N" + std::to_string(N[d]);
    return s;
}

// build_solution_filename (dim_output_utils.hpp:49-75)
template<int DIM>
std::string solution_filename(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, const std::string& suffix)
{
    std::string f = c.output_prefix;
    if (!c.output_dir.empty()) f = c.output_dir + "/" + f + "/" + f;
    if (!suffix.empty()) f += "_" + suffix;
    return f + grid_suffix<DIM>(N) + ".csv";
}

// <output_dir>/<prefix>/<prefix>_N.._<tail> (runtime_report.hpp:14-50,
// conservation_report.hpp:53-88)
template<int DIM>
std::filesystem::path report_filename(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, const std::string& tail)
{
    const std::string stem = c.output_prefix + grid_suffix<DIM>(N) + tail;
    if (c.output_dir.empty()) return std::filesystem::path(stem);
    return std::filesystem::path(c.output_dir) / c.output_prefix / stem;
}

inline void make_parent(const std::filesystem::path& p)
{
    if (!p.parent_path().empty()) std::filesystem::create_directories(p.parent_path());
}

// write_csv (dim_write_csv.hpp:14-124).  rows: per cell
// rho, vel[DIM], p, e, alpha[nmat], rho_mat[nmat], mass[nmat].
template<int DIM>
void write_solution_csv(const std::string& filename, const CaseConfig<DIM>& c, const std::array<int, DIM>& N, int nmat,
                        const std::vector<double>& rows, unsigned threads = 0)
{
    std::cout << "Writing to: " << filename << "\n";
    make_parent(std::filesystem::path(filename));
    std::ofstream file(filename, std::ios::binary);
    if (!file) throw std::runtime_error("dim::write_csv: failed to open file");
    std::array<double, DIM> dx{};
    for (int d = 0; d < DIM; ++d) dx[d] = (c.domain_max[d] - c.domain_min[d]) / static_cast<double>(N[d]);
    std::string head;
    for (int d = 0; d < DIM; ++d) head += "x" + std::to_string(d) + ",";
    head += "rho,";
    for (int d = 0; d < DIM; ++d) head += "u" + std::to_string(d) + ",";
    head += "p,e";
    for (int k = 0; k < nmat; ++k) head += ",alpha" + std::to_string(k);
    for (int k = 0; k < nmat; ++k) head += ",rho_mat" + std::to_string(k);
    for (int k = 0; k < nmat; ++k) head += ",mass" + std::to_string(k);
    head += "\n";
    file << head;

    const int nr = 3 + DIM + 3 * nmat;
    long long total = 1;
    for (int d = 0; d < DIM; ++d) total *= N[d];
    if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    const long long block = 1 << 16;
    std::vector<std::string> texts(threads);
    for (long long first = 0; first < total; first += block * threads) {
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < threads; ++t) {
            pool.emplace_back([&, t] {
                std::string& s = texts[t];
                s.clear();
                const long long beg = first + t * block, end = std::min(total, beg + block);
                for (long long i = beg; i < end; ++i) {
                    long long stride = 1;
                    for (int d = 0; d < DIM; ++d) {
                        const int idx = (int)((i / stride) % N[d]);
                        stride *= N[d];
                        put_number(s, c.domain_min[d] + (idx + 0.5) * dx[d]);
                        s += ',';
                    }
                    const double* r = &rows[(size_t)i * nr];
                    int o = 0;
                    put_number(s, r[o++]);  // rho
                    s += ',';
                    for (int d = 0; d < DIM; ++d) {
                        put_number(s, r[o++]);
                        s += ',';
                    }
                    put_number(s, r[o++]);  // p
                    s += ',';
                    put_number(s, r[o++]);  // e
                    for (int k = 0; k < 3 * nmat; ++k) {
                        s += ',';
                        put_number(s, r[o++]);
                    }
                    s += '\n';
                }
            });
        }
        for (auto& th : pool) th.join();
        for (const auto& s : texts) file << s;
    }
    if (!file) throw std::runtime_error("dim::write_csv: failed to open file");
    std::cout << "Written: " << filename << "\n";
}

// SolverPhaseTimings (src/core/phase_timings.hpp:4-21)
struct PhaseTimes {
    double cfl = 0.0, sweep = 0.0, boundary = 0.0, conservation = 0.0, output = 0.0;
};

inline std::string env_or_empty(const char* name)
{
    const char* v = std::getenv(name);
    return v ? v : "";
}

inline std::string cpus_allowed(const std::string& status_path)
{
    std::ifstream in(status_path);
    std::string line;
    const std::string key = "Cpus_allowed_list:";
    while (std::getline(in, line)) {
        if (line.rfind(key, 0) == 0) {
            const auto v = line.substr(key.size());
            const auto a = v.find_first_not_of(" \t");
            return a == std::string::npos ? "" : v.substr(a);
        }
    }
    return "";
}

// write_runtime_report (runtime_report.hpp:64-149) with the keys of the
// reference; the OpenMP lines describe this host process (the time loop runs
// on the GPU, one host thread), and two lines name the GPU form.
template<int DIM>
void write_runtime_report(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, long long steps, double final_time,
                          double wall_seconds, const PhaseTimes& ph, const std::string& precision)
{
    const auto path = report_filename<DIM>(c, N, "_runtime.txt");
    make_parent(path);
    std::ofstream f(path);
    if (!f) throw std::runtime_error("write_runtime_report: failed to open " + path.string());
    long long cells = 1;
    for (int d = 0; d < DIM; ++d) cells *= N[d];
    const double updates = (double)cells * (double)steps;
    f << "interface_method = " << c.interface_method << "\n";
    f << "dimension = " << DIM << "\n";
    f << "N = ";
    for (int d = 0; d < DIM; ++d) f << (d ? "x" : "") << N[d];
    f << "\n";
    f << "cells = " << cells << "\n";
    f << "steps = " << steps << "\n";
    f << "final_time = " << final_time << "\n";
    f << "wall_time_seconds = " << wall_seconds << "\n";
    f << "cost_per_cell_update_seconds = " << (updates > 0.0 ? wall_seconds / updates : 0.0) << "\n";
    const std::string allowed = cpus_allowed("/proc/self/status");
    f << "openmp_compiled = false\n";
    f << "openmp_num_threads_env = " << env_or_empty("OMP_NUM_THREADS") << "\n";
    f << "openmp_max_threads = 1\n";
    f << "openmp_observed_threads = 1\n";
    f << "openmp_dynamic_enabled = 0\n";
    f << "openmp_proc_bind_env = " << env_or_empty("OMP_PROC_BIND") << "\n";
    f << "openmp_places_env = " << env_or_empty("OMP_PLACES") << "\n";
    f << "openmp_schedule_env = " << env_or_empty("OMP_SCHEDULE") << "\n";
    f << "quant_cpu_list_env = " << env_or_empty("QUANT_CPU_LIST") << "\n";
    f << "cpus_allowed_list = " << allowed << "\n";
    f << "openmp_thread_cpus_allowed_lists = " << allowed << "\n";
    f << "phase_cfl_seconds = " << ph.cfl << "\n";
    f << "phase_sweep_seconds = " << ph.sweep << "\n";
    f << "phase_boundary_seconds = " << ph.boundary << "\n";
    f << "phase_level_set_seconds = " << 0.0 << "\n";
    f << "phase_material_transfer_seconds = " << 0.0 << "\n";
    f << "phase_conservation_seconds = " << ph.conservation << "\n";
    f << "phase_output_seconds = " << ph.output << "\n";
    f << "solver_backend = dimhip (gfx950)\n";
    f << "precision = " << precision << "\n";
}

// Totals in the library's order: [mass_k (nmat), mass, mom_d (DIM), E].
struct Totals {
    std::vector<double> v;
    int nmat = 0, dim = 0;
    double mass() const { return v[nmat]; }
    double mom(int d) const { return v[nmat + 1 + d]; }
    double energy() const { return v[nmat + 1 + dim]; }
    double material(int k) const { return v[k]; }
};

// ConservationReport (conservation_report.hpp:477-698) for the DIM path: the
// interface flux mismatch integral is zero for DIM (run_dim.hpp:113-125
// passes none).
class ConservationLog {
public:
    // path: report_filename(cfg, N, "_conservation.csv")
    ConservationLog(const std::filesystem::path& path, const Totals& initial) : initial_(initial), scale_(initial)
    {
        make_parent(path);
        file_.open(path);
        if (!file_) throw std::runtime_error("ConservationReport: failed to open " + path.string());
        const int M = initial.nmat, D = initial.dim;
        file_ << "step,time,mass";
        for (int d = 0; d < D; ++d) file_ << ",momentum" << d;
        file_ << ",energy,eps_mass";
        for (int d = 0; d < D; ++d) file_ << ",eps_momentum" << d;
        file_ << ",eps_energy";
        for (int k = 0; k < M; ++k) file_ << ",material_mass" << k << ",eps_material_mass" << k;
        balance_header("mass");
        for (int d = 0; d < D; ++d) balance_header("momentum" + std::to_string(d));
        balance_header("energy");
        for (int k = 0; k < M; ++k) balance_header("material_mass" + std::to_string(k));
        file_ << "\n";
    }

    // flux_integral: the time integral of the boundary flux, library order.
    void write(long long step, double time, const Totals& cur, const std::vector<double>& flux_integral)
    {
        const int M = cur.nmat, D = cur.dim;
        for (size_t i = 0; i < cur.v.size(); ++i) scale_.v[i] = std::max(scale_.v[i], std::abs(cur.v[i]));
        file_ << step << "," << time << "," << cur.mass();
        for (int d = 0; d < D; ++d) file_ << "," << cur.mom(d);
        file_ << "," << cur.energy() << "," << rel(cur.mass(), initial_.mass());
        for (int d = 0; d < D; ++d) file_ << "," << rel(cur.mom(d), initial_.mom(d));
        file_ << "," << rel(cur.energy(), initial_.energy());
        for (int k = 0; k < M; ++k) file_ << "," << cur.material(k) << "," << rel(cur.material(k), initial_.material(k));
        const auto col = [&](int i) { balance(cur.v[i], initial_.v[i], flux_integral[i], scale_.v[i]); };
        col(M);
        for (int d = 0; d < D; ++d) col(M + 1 + d);
        col(M + 1 + D);
        for (int k = 0; k < M; ++k) col(k);
        file_ << "\n";
    }

private:
    std::ofstream file_;
    Totals initial_, scale_;

    static double rel(double v, double v0, double floor = 1.0e-12)
    {
        return std::abs(v - v0) / std::max(std::abs(v0), floor);
    }

    void balance_header(const std::string& v)
    {
        file_ << ",raw_drift_" << v << ",boundary_flux_" << v << ",balance_residual_" << v
              << ",normalized_balance_residual_" << v << ",interface_flux_mismatch_" << v << ",near_zero_initial_" << v;
    }

    void balance(double cur, double init, double flux, double scale)
    {
        const double floor = 1.0e-12;
        const double drift = cur - init;
        const double residual = drift + flux;
        const double denom = std::max({std::abs(init), std::abs(scale), floor});
        const int near_zero = std::abs(init) <= std::max(floor, 1.0e-10 * std::max(std::abs(scale), floor)) ? 1 : 0;
        file_ << "," << drift << "," << flux << "," << residual << "," << std::abs(residual) / denom << "," << 0.0
              << "," << near_zero;
    }
};

}  // namespace dimhost
