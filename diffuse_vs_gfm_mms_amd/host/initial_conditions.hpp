// initial_conditions.hpp — the DIM initial states of the reference's case
// files, as primitive rows rho[nmat], alpha[nmat], vel[DIM], p per cell.
//
// The geometry and blending follow src/setup/dim_initial_conditions.hpp:16-628
// operation for operation: region boxes and oblique planar slabs with tanh
// blending, the shock-bubble and the coated shock-bubble.  The host builds the
// primitive rows (std::tanh of the host C library, the reference's own); the
// library converts them with prim_to_cons on the device
// (dimhip_upload_primitive), so the initial state is the reference's bit for
// bit.  Rows are built in parallel threads; each cell's arithmetic is
// independent of the thread count.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <thread>
#include <vector>

#include "config.hpp"

namespace dimhost {

// One cell's primitive (the reference's dim::Primitive with nmat <= 4).
template<int DIM>
struct CellPrim {
    int nmat = 0;
    double rho[4] = {1.0, 1.0, 1.0, 1.0};
    double alpha[4] = {0.0, 0.0, 0.0, 0.0};
    std::array<double, DIM> vel{};
    double p = 0.0;
};

// sanitise_alpha (src/dim/primitives.hpp:45-70)
inline void sanitise_fractions(double* alpha, int n, double floor)
{
    double sum = 0.0;
    for (int k = 0; k < n; ++k) {
        double v = std::isfinite(alpha[k]) ? alpha[k] : floor;
        v = (v < floor) ? floor : v;
        alpha[k] = v;
        sum += v;
    }
    if (!std::isfinite(sum) || sum <= 0.0) {
        for (int k = 0; k < n; ++k) alpha[k] = 0.0;
        alpha[0] = 1.0;
        return;
    }
    for (int k = 0; k < n; ++k) alpha[k] /= sum;
}

template<int DIM>
class InitialState {
public:
    InitialState(const CaseConfig<DIM>& c, const std::array<int, DIM>& N, int nmat) : c_(c), N_(N), nmat_(nmat)
    {
        for (int d = 0; d < DIM; ++d) dx_[d] = (c.domain_max[d] - c.domain_min[d]) / static_cast<double>(N[d]);
        planar_ = c.initial_condition == "planar_regions";
        if (planar_) {
            // normalised_planar_normal (dim_initial_conditions.hpp:33-55)
            double n2 = 0.0;
            for (int d = 0; d < DIM; ++d) n2 += c.planar_normal[d] * c.planar_normal[d];
            if (n2 <= 0.0) throw std::runtime_error("dim::normalised_planar_normal: zero planar normal");
            const double inv = 1.0 / std::sqrt(n2);
            for (int d = 0; d < DIM; ++d) normal_[d] = c.planar_normal[d] * inv;
        }
        const std::string& ic = c.initial_condition;
        if (ic == "regions" || ic == "planar_regions") kind_ = 0;
        else if (ic == "shock_bubble") kind_ = 1;
        else if (ic == "coated_shock_bubble") kind_ = 2;
        else throw std::runtime_error("dim::initialise_dim_from_config: unsupported initial_condition: " + ic);
    }

    // All cells, x-fastest: rows of 2*nmat + DIM + 1 doubles.
    std::vector<double> rows(unsigned threads = 0) const
    {
        long long total = 1;
        for (int d = 0; d < DIM; ++d) {
            if (N_[d] <= 0) throw std::runtime_error("dim::total_cells: grid size must be positive");
            total *= N_[d];
        }
        const int np = 2 * nmat_ + DIM + 1;
        std::vector<double> out((size_t)total * np);
        if (threads == 0) threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        std::vector<std::exception_ptr> errors(threads);
        const long long chunk = (total + threads - 1) / threads;
        for (unsigned t = 0; t < threads; ++t) {
            pool.emplace_back([&, t] {
                try {
                    const long long beg = t * chunk, end = std::min(total, beg + chunk);
                    for (long long i = beg; i < end; ++i) write_row(i, &out[(size_t)i * np]);
                } catch (...) {
                    errors[t] = std::current_exception();
                }
            });
        }
        for (auto& th : pool) th.join();
        for (auto& e : errors)
            if (e) std::rethrow_exception(e);
        return out;
    }

    CellPrim<DIM> at(const std::array<double, DIM>& x) const
    {
        if (kind_ == 0) return diffuse_region(x);
        if (kind_ == 1) return shock_bubble(x);
        return coated_shock_bubble(x);
    }

private:
    const CaseConfig<DIM>& c_;
    std::array<int, DIM> N_;
    int nmat_;
    std::array<double, DIM> dx_{};
    std::array<double, DIM> normal_{};
    bool planar_ = false;
    int kind_ = 0;

    void write_row(long long linear, double* row) const
    {
        // unflatten_index + compute_cell_center (geometry.hpp:57-89)
        std::array<double, DIM> x{};
        long long stride = 1;
        for (int d = 0; d < DIM; ++d) {
            const int idx = (int)((linear / stride) % N_[d]);
            stride *= N_[d];
            x[d] = c_.domain_min[d] + (idx + 0.5) * dx_[d];
        }
        const CellPrim<DIM> P = at(x);
        int o = 0;
        for (int k = 0; k < nmat_; ++k) row[o++] = P.rho[k];
        for (int k = 0; k < nmat_; ++k) row[o++] = P.alpha[k];
        for (int d = 0; d < DIM; ++d) row[o++] = P.vel[d];
        row[o++] = P.p;
    }

    CellPrim<DIM> blank() const
    {
        CellPrim<DIM> P;
        P.nmat = nmat_;
        return P;
    }

    int slot(int material_id) const
    {
        for (int k = 0; k < (int)c_.materials.size(); ++k)
            if (c_.materials[k].id == material_id) return k;
        throw std::runtime_error("dim::material_slot_from_id: material id not found");
    }

    double planar_coordinate(const std::array<double, DIM>& x) const
    {
        double s = 0.0;
        for (int d = 0; d < DIM; ++d) s += normal_[d] * x[d];
        return s;
    }

    bool inside(const std::array<double, DIM>& x, const RegionSpec<DIM>& r) const
    {
        if (planar_) {
            const double s = planar_coordinate(x);
            return s >= r.lower[0] && s < r.upper[0];
        }
        for (int d = 0; d < DIM; ++d)
            if (x[d] < r.lower[d] || x[d] >= r.upper[d]) return false;
        return true;
    }

    // signed_distance_to_box (dim_initial_conditions.hpp:81-133): axes the box
    // spans from domain edge to domain edge carry no interface.
    double box_distance(const std::array<double, DIM>& x, const RegionSpec<DIM>& r) const
    {
        double outside_sq = 0.0;
        double inside_dist = std::numeric_limits<double>::max();
        bool in = true, any_axis = false;
        for (int d = 0; d < DIM; ++d) {
            const bool lo_edge = std::abs(r.lower[d] - c_.domain_min[d]) <= 1e-14;
            const bool hi_edge = std::abs(r.upper[d] - c_.domain_max[d]) <= 1e-14;
            if (lo_edge && hi_edge) continue;
            any_axis = true;
            if (x[d] < r.lower[d]) {
                const double gap = r.lower[d] - x[d];
                outside_sq += gap * gap;
                in = false;
            } else if (x[d] > r.upper[d]) {
                const double gap = x[d] - r.upper[d];
                outside_sq += gap * gap;
                in = false;
            } else {
                const double a = x[d] - r.lower[d];
                const double b = r.upper[d] - x[d];
                inside_dist = std::min(inside_dist, std::min(a, b));
            }
        }
        if (!any_axis) return -std::numeric_limits<double>::max();
        return in ? -inside_dist : std::sqrt(outside_sq);
    }

    // signed_distance_to_planar_interval (dim_initial_conditions.hpp:150-167)
    double slab_distance(const std::array<double, DIM>& x, const RegionSpec<DIM>& r) const
    {
        const double s = planar_coordinate(x);
        if (s < r.lower[0]) return r.lower[0] - s;
        if (s > r.upper[0]) return s - r.upper[0];
        return -std::min(s - r.lower[0], r.upper[0] - s);
    }

    static double sphere_distance(const std::array<double, DIM>& x, const std::array<double, DIM>& centre, double radius)
    {
        double r2 = 0.0;
        for (int d = 0; d < DIM; ++d) {
            const double delta = x[d] - centre[d];
            r2 += delta * delta;
        }
        return std::sqrt(r2) - radius;
    }

    // smooth_indicator (dim_initial_conditions.hpp:218-229)
    double weight(double signed_distance) const
    {
        if (c_.interface_thickness <= 0.0) return (signed_distance <= 0.0) ? 1.0 : 0.0;
        const double a = std::max(c_.interface_sharpness_alpha, 1.0e-12);
        return 0.5 * (1.0 - std::tanh(a * signed_distance / c_.interface_thickness));
    }

    // sharp_region_primitive (dim_initial_conditions.hpp:185-216)
    CellPrim<DIM> sharp_region(const std::array<double, DIM>& x) const
    {
        for (const auto& r : c_.regions) {
            if (!inside(x, r)) continue;
            CellPrim<DIM> P = blank();
            P.vel = r.vel;
            P.p = r.p;
            const int s = slot(r.material_id);
            P.alpha[s] = 1.0;
            P.rho[s] = r.rho;
            return P;
        }
        throw std::runtime_error("dim::sharp_region_primitive: point not covered by any region");
    }

    // diffuse_region_primitive (dim_initial_conditions.hpp:232-312)
    CellPrim<DIM> diffuse_region(const std::array<double, DIM>& x) const
    {
        if (c_.regions.empty()) throw std::runtime_error("dim::diffuse_region_primitive: no regions defined");
        if (c_.interface_thickness <= 0.0) return sharp_region(x);
        double score_k[4] = {0, 0, 0, 0}, rho_k[4] = {0, 0, 0, 0};
        double total = 0.0, p_sum = 0.0;
        std::array<double, DIM> vel_sum{};
        for (const auto& r : c_.regions) {
            const double sd = planar_ ? slab_distance(x, r) : box_distance(x, r);
            const double w = weight(sd);
            if (w <= 1e-14) continue;
            const int s = slot(r.material_id);
            score_k[s] += w;
            rho_k[s] += w * r.rho;
            p_sum += w * r.p;
            for (int d = 0; d < DIM; ++d) vel_sum[d] += w * r.vel[d];
            total += w;
        }
        if (total <= 1e-14) return sharp_region(x);
        CellPrim<DIM> P = blank();
        P.p = p_sum / total;
        for (int d = 0; d < DIM; ++d) P.vel[d] = vel_sum[d] / total;
        double alpha_sum = 0.0;
        for (int k = 0; k < nmat_; ++k) {
            P.alpha[k] = score_k[k] / total;
            alpha_sum += P.alpha[k];
            if (score_k[k] > 1e-14) P.rho[k] = rho_k[k] / score_k[k];
        }
        if (alpha_sum <= 0.0) return sharp_region(x);
        sanitise_fractions(P.alpha, nmat_, 0.0);
        return P;
    }

    // Adds one material's blending weight (the add_material lambdas of
    // dim_initial_conditions.hpp:364-374,487-497).
    static void blend(CellPrim<DIM>& P, double* rho_sum, double* w_sum, int s, double w, double rho)
    {
        if (w <= 1e-14) return;
        P.alpha[s] += w;
        rho_sum[s] += w * rho;
        w_sum[s] += w;
    }

    // diffuse_shock_bubble_primitive (dim_initial_conditions.hpp:315-403)
    CellPrim<DIM> shock_bubble(const std::array<double, DIM>& x) const
    {
        const double sd = sphere_distance(x, c_.bubble_center, c_.bubble_radius);
        const PhaseState<DIM>& gas = x[c_.shock_axis] >= c_.shock_position ? c_.right : c_.left;
        const int gas_slot = slot(gas.material);
        const int bubble_slot = slot(c_.bubble.material);
        CellPrim<DIM> P = blank();
        if (c_.interface_thickness <= 0.0) {
            const bool in = sd <= 0.0;
            const int s = in ? bubble_slot : gas_slot;
            P.alpha[s] = 1.0;
            P.rho[s] = in ? c_.bubble.rho : gas.rho;
            P.vel = in ? c_.bubble.vel : gas.vel;
            P.p = in ? c_.bubble.p : gas.p;
            return P;
        }
        const double wb = weight(sd);
        const double wg = 1.0 - wb;
        double rho_sum[4] = {0, 0, 0, 0}, w_sum[4] = {0, 0, 0, 0};
        blend(P, rho_sum, w_sum, gas_slot, wg, gas.rho);
        blend(P, rho_sum, w_sum, bubble_slot, wb, c_.bubble.rho);
        for (int k = 0; k < nmat_; ++k)
            if (w_sum[k] > 1e-14) P.rho[k] = rho_sum[k] / w_sum[k];
        for (int d = 0; d < DIM; ++d) P.vel[d] = wg * gas.vel[d] + wb * c_.bubble.vel[d];
        P.p = wg * gas.p + wb * c_.bubble.p;
        sanitise_fractions(P.alpha, nmat_, 0.0);
        return P;
    }

    // diffuse_coated_shock_bubble_primitive (dim_initial_conditions.hpp:406-522)
    CellPrim<DIM> coated_shock_bubble(const std::array<double, DIM>& x) const
    {
        const double sd_in = sphere_distance(x, c_.bubble_center, c_.bubble_radius);
        const double sd_out = sphere_distance(x, c_.bubble_center, c_.film_radius);
        const PhaseState<DIM>& gas = x[c_.shock_axis] >= c_.shock_position ? c_.right : c_.left;
        const int gas_slot = slot(gas.material);
        const int film_slot = slot(c_.film.material);
        const int bubble_slot = slot(c_.bubble.material);
        CellPrim<DIM> P = blank();
        if (c_.interface_thickness <= 0.0) {
            int s = gas_slot;
            if (sd_in <= 0.0) {
                s = bubble_slot;
                P.rho[s] = c_.bubble.rho;
                P.vel = c_.bubble.vel;
                P.p = c_.bubble.p;
            } else if (sd_out <= 0.0) {
                s = film_slot;
                P.rho[s] = c_.film.rho;
                P.vel = c_.film.vel;
                P.p = c_.film.p;
            } else {
                P.rho[s] = gas.rho;
                P.vel = gas.vel;
                P.p = gas.p;
            }
            P.alpha[s] = 1.0;
            return P;
        }
        const double wb = weight(sd_in);
        const double wc = weight(sd_out);
        const double wf = std::max(wc - wb, 0.0);
        const double wg = std::max(1.0 - wc, 0.0);
        double rho_sum[4] = {0, 0, 0, 0}, w_sum[4] = {0, 0, 0, 0};
        blend(P, rho_sum, w_sum, gas_slot, wg, gas.rho);
        blend(P, rho_sum, w_sum, film_slot, wf, c_.film.rho);
        blend(P, rho_sum, w_sum, bubble_slot, wb, c_.bubble.rho);
        const double wt = std::max(wg + wf + wb, 1e-14);
        for (int k = 0; k < nmat_; ++k)
            if (w_sum[k] > 1e-14) P.rho[k] = rho_sum[k] / w_sum[k];
        for (int d = 0; d < DIM; ++d) P.vel[d] = (wg * gas.vel[d] + wf * c_.film.vel[d] + wb * c_.bubble.vel[d]) / wt;
        P.p = (wg * gas.p + wf * c_.film.p + wb * c_.bubble.p) / wt;
        sanitise_fractions(P.alpha, nmat_, 0.0);
        return P;
    }
};

}  // namespace dimhost
