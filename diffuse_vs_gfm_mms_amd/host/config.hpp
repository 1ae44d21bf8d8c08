// config.hpp — the reference's `key = value` case file, read for the DIM path.
//
// Format and validation follow src/io/config_loader.hpp:18-570 and the
// defaults of src/io/config.hpp:34-136: '#' starts a comment, one key per line,
// arrays in [ ], material and region lines as comma lists with key=value
// options.  Every key the reference accepts is accepted here (a case file
// written for the reference loads unchanged); keys that only steer the
// sharp-interface, level-set or solid solvers are kept but unused.  Error texts
// are the reference's, so `ERROR:` lines match.
#pragma once

#include <algorithm>
#include <array>
#include <cctype>
#include <fstream>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace dimhost {

struct MaterialSpec {
    int id = -1;
    std::string type;
    std::map<std::string, double> params;
};

template<int DIM>
struct RegionSpec {
    std::array<double, DIM> lower{}, upper{};
    double rho = 0.0;
    std::array<double, DIM> vel{};
    double p = 0.0;
    int material_id = -1;
};

// A uniform state of one material (the *_left, *_right, *_bubble, *_film,
// *_in, *_out key groups).
template<int DIM>
struct PhaseState {
    double rho = 0.0;
    std::array<double, DIM> vel{};
    double p = 0.0;
    int material = -1;
};

template<int DIM>
struct CaseConfig {
    std::string model_type = "auto";
    std::array<double, DIM> domain_min{}, domain_max{};
    std::vector<std::array<int, DIM>> N_list;
    double tfinal = 0.0;
    std::vector<double> output_times;
    double cfl = 0.0;
    bool exact_riemann = false;
    std::string output_prefix = "output";
    std::string output_dir = "output";
    std::vector<MaterialSpec> materials;
    std::vector<RegionSpec<DIM>> regions;
    std::string initial_condition = "regions";
    std::array<double, DIM> planar_normal{};
    std::array<double, DIM> explosion_center{};
    std::vector<std::array<double, DIM>> explosion_centers;
    double explosion_radius = 0.0;
    PhaseState<DIM> in, out, left, right, bubble, film;
    int shock_axis = 0;
    double shock_position = 0.0;
    std::array<double, DIM> bubble_center{};
    double bubble_radius = 0.0;
    double film_radius = 0.0;
    std::string interface_method = "SM";
    std::string time_update = "split";
    bool use_level_set = false;
    int reinit_interval = 0;
    int reinit_iterations = 10;
    std::string level_set_reinit_method = "sussman";
    std::string level_set_advection = "normal_speed";
    std::string level_set_spatial_derivative = "weno2";
    std::string level_set_component_policy = "largest_overlap";
    bool rgfm_diagnostics = false;
    int rgfm_diagnostics_interval = 1;
    std::string rgfm_star_velocity_mode = "input_mean";
    double interface_thickness = 0.0;
    double interface_sharpness_alpha = 2.0;
    double dim_alpha_source_floor = 0.0;
    std::string dim_lambda_model = "kapila";
    int barton_solid_material = -1;
    double barton_temperature = 300.0;
    std::array<std::string, DIM> bc_lo, bc_hi;

    CaseConfig()
    {
        bc_lo.fill("transmissive");
        bc_hi.fill("transmissive");
    }
};

namespace text {

inline std::string trim(const std::string& s)
{
    const char* ws = " \t\r\n";
    const auto a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

inline std::string lower(std::string s)
{
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

// Split on `delim` outside [ ] nesting; an unbalanced bracket is an error.
inline std::vector<std::string> split_outside_brackets(const std::string& s, char delim)
{
    std::vector<std::string> parts(1);
    int depth = 0;
    for (char ch : s) {
        depth += (ch == '[') - (ch == ']');
        if (ch == delim && depth == 0) parts.emplace_back();
        else parts.back().push_back(ch);
    }
    if (depth != 0) throw std::runtime_error("Unmatched brackets in: " + s);
    for (auto& p : parts) p = trim(p);
    return parts;
}

inline std::pair<std::string, std::string> key_value(const std::string& s)
{
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        depth += (s[i] == '[') - (s[i] == ']');
        if (s[i] == '=' && depth == 0) return {trim(s.substr(0, i)), trim(s.substr(i + 1))};
    }
    throw std::runtime_error("Invalid key=value: " + s);
}

inline std::string bracket_body(const std::string& s)
{
    const std::string t = trim(s);
    if (t.empty() || t.front() != '[' || t.back() != ']') throw std::runtime_error("Expected [ ... ]: " + s);
    return trim(t.substr(1, t.size() - 2));
}

inline bool boolean(const std::string& s)
{
    const std::string v = lower(trim(s));
    if (v == "true" || v == "1" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "no") return false;
    throw std::runtime_error("Invalid boolean: " + s);
}

template<int DIM, typename T, typename Conv>
std::array<T, DIM> fixed_array(const std::string& s, Conv conv, const char* err)
{
    const auto parts = split_outside_brackets(bracket_body(s), ',');
    if ((int)parts.size() != DIM) throw std::runtime_error(err);
    std::array<T, DIM> out{};
    for (int i = 0; i < DIM; ++i) out[i] = conv(parts[i]);
    return out;
}

template<int DIM>
std::array<double, DIM> doubles(const std::string& s)
{
    return fixed_array<DIM, double>(s, [](const std::string& t) { return std::stod(t); }, "Wrong array size");
}

template<int DIM>
std::array<int, DIM> ints(const std::string& s)
{
    return fixed_array<DIM, int>(s, [](const std::string& t) { return std::stoi(t); }, "Wrong int array size");
}

template<int DIM>
std::array<std::string, DIM> names(const std::string& s)
{
    return fixed_array<DIM, std::string>(s, [](const std::string& t) { return lower(trim(t)); },
                                         "Wrong string array size");
}

inline std::vector<double> double_list(const std::string& s)
{
    const std::string body = bracket_body(s);
    std::vector<double> out;
    if (trim(body).empty()) return out;
    for (const auto& t : split_outside_brackets(body, ',')) out.push_back(std::stod(t));
    return out;
}

template<int DIM>
std::vector<std::array<double, DIM>> array_list(const std::string& s)
{
    std::vector<std::array<double, DIM>> out;
    for (const auto& t : split_outside_brackets(bracket_body(s), ',')) out.push_back(doubles<DIM>(t));
    return out;
}

inline std::string model_name(std::string v)
{
    v = trim(v);
    for (char& ch : v) ch = (ch == '-' || ch == ' ') ? '_' : (char)std::tolower((unsigned char)ch);
    return v;
}

}  // namespace text

// Model family of a model_type / interface_method name (src/core/model_type.hpp:36-101).
inline std::string model_family(const std::string& raw, bool from_method)
{
    const std::string v = text::model_name(raw);
    if (from_method) {
        if (v == "sm") return "euler";
        if (v == "dim") return "dim";
        if (v == "gfm" || v == "sim" || v == "rgfm") return "sharp";
        throw std::runtime_error("Cannot infer model_type from interface_method: " + raw);
    }
    if (v.empty() || v == "auto") return "auto";
    if (v == "sm" || v == "single_material" || v == "single_material_euler" || v == "euler") return "euler";
    if (v == "dim" || v == "diffuse_interface" || v == "diffuse_interface_model") return "dim";
    if (v == "sim" || v == "gfm" || v == "rgfm" || v == "sharp_interface" || v == "ghost_fluid") return "sharp";
    if (v == "elastic_solid" || v == "solid_elastic") return "elastic";
    if (v == "elastoplastic_solid" || v == "solid_elastoplastic" || v == "elasto_plastic_solid" ||
        v == "plastic_solid")
        return "elastoplastic";
    throw std::runtime_error("Unknown model_type: " + raw);
}

inline MaterialSpec parse_material(const std::string& value)
{
    const auto parts = text::split_outside_brackets(value, ',');
    if (parts.size() < 2) throw std::runtime_error("Invalid material line");
    MaterialSpec m;
    m.id = std::stoi(parts[0]);
    m.type = parts[1];
    for (size_t i = 2; i < parts.size(); ++i) {
        const auto kv = text::key_value(parts[i]);
        m.params[kv.first] = std::stod(kv.second);
    }
    return m;
}

template<int DIM>
RegionSpec<DIM> parse_region(const std::string& value)
{
    const auto parts = text::split_outside_brackets(value, ',');
    if (parts.size() < 2) throw std::runtime_error("Invalid region line");
    RegionSpec<DIM> r;
    r.lower = text::doubles<DIM>(parts[0]);
    r.upper = text::doubles<DIM>(parts[1]);
    for (size_t i = 2; i < parts.size(); ++i) {
        const auto kv = text::key_value(parts[i]);
        if (kv.first == "rho") r.rho = std::stod(kv.second);
        else if (kv.first == "p") r.p = std::stod(kv.second);
        else if (kv.first == "vel") r.vel = text::doubles<DIM>(kv.second);
        else if (kv.first == "material") r.material_id = std::stoi(kv.second);
        else throw std::runtime_error("Unknown region key: " + kv.first);
    }
    return r;
}

// The checks of validate_config (config_loader.hpp:237-438), in its order.
template<int DIM>
void validate(const CaseConfig<DIM>& c)
{
    if (c.materials.empty()) throw std::runtime_error("No materials defined");
    std::vector<bool> seen(c.materials.size(), false);
    for (const auto& m : c.materials) {
        if (m.id < 0 || m.id >= (int)c.materials.size())
            throw std::runtime_error("material ids must be contiguous and zero-based: expected ids 0.." +
                                     std::to_string(c.materials.size() - 1));
        if (seen[m.id]) throw std::runtime_error("duplicate material id: " + std::to_string(m.id));
        seen[m.id] = true;
    }
    for (size_t id = 0; id < seen.size(); ++id)
        if (!seen[id]) throw std::runtime_error("missing material id in contiguous material set: " + std::to_string(id));
    if (c.N_list.empty()) throw std::runtime_error("No grid resolution specified");
    for (double t : c.output_times)
        if (t < 0.0) throw std::runtime_error("output_times must be non-negative");
    if (!std::is_sorted(c.output_times.begin(), c.output_times.end()))
        throw std::runtime_error("output_times must be sorted in ascending order");
    for (size_t i = 1; i < c.output_times.size(); ++i)
        if (c.output_times[i] <= c.output_times[i - 1]) throw std::runtime_error("output_times must be strictly increasing");
    if (!c.output_times.empty()) {
        if (c.tfinal <= 0.0) throw std::runtime_error("tfinal must be positive when output_times is set");
        if (c.output_times.back() > c.tfinal + 1e-14) throw std::runtime_error("output_times cannot exceed tfinal");
    }
    if (c.interface_method != "SM" && c.interface_method != "GFM" && c.interface_method != "DIM")
        throw std::runtime_error("Invalid interface_method");
    const std::string fam = model_family(c.model_type, false);
    if (fam != "auto" && !(fam == "elastoplastic" && text::model_name(c.interface_method) == "dim") &&
        fam != model_family(c.interface_method, true))
        throw std::runtime_error("model_type does not match interface_method");
    if (c.time_update != "split" && c.time_update != "unsplit")
        throw std::runtime_error("time_update must be split or unsplit");
    if (c.interface_method == "GFM" && !c.use_level_set) throw std::runtime_error("GFM requires level set");
    const auto one_of = [](const std::string& v, std::initializer_list<const char*> allowed) {
        for (const char* a : allowed)
            if (v == a) return true;
        return false;
    };
    if (!one_of(c.level_set_advection, {"normal_speed", "rgfm_normal_speed", "flow", "physical_flow"}))
        throw std::runtime_error("level_set_advection must be normal_speed, rgfm_normal_speed, flow, or physical_flow");
    if (!one_of(c.level_set_spatial_derivative, {"weno2", "first_order"}))
        throw std::runtime_error("level_set_spatial_derivative must be weno2 or first_order");
    if (!one_of(c.level_set_component_policy, {"largest_overlap", "all_overlapping"}))
        throw std::runtime_error("level_set_component_policy must be largest_overlap or all_overlapping");
    if (c.reinit_interval < 0) throw std::runtime_error("reinit_interval must be non-negative");
    if (c.reinit_iterations < 0) throw std::runtime_error("reinit_iterations must be non-negative");
    if (!one_of(c.level_set_reinit_method, {"sussman", "redistance"}))
        throw std::runtime_error("level_set_reinit_method must be sussman or redistance");
    if (c.rgfm_diagnostics_interval <= 0) throw std::runtime_error("rgfm_diagnostics_interval must be positive");
    if (!one_of(c.rgfm_star_velocity_mode, {"mcrs", "input_mean", "zero"}))
        throw std::runtime_error("rgfm_star_velocity_mode must be mcrs, input_mean, or zero");
    if (c.interface_method == "DIM" && c.use_level_set) throw std::runtime_error("DIM should not use level set");
    for (int d = 0; d < DIM; ++d)
        for (const auto* bc : {&c.bc_lo[d], &c.bc_hi[d]})
            if (!one_of(*bc, {"transmissive", "zero_gradient", "outflow", "nonreflective", "non_reflective",
                              "reflective", "reflection"}))
                throw std::runtime_error("Invalid boundary condition");
    const bool region_ic = c.initial_condition == "regions" || c.initial_condition == "planar_regions";
    if (region_ic && c.regions.empty()) throw std::runtime_error("No regions defined");
    if (c.initial_condition == "planar_regions") {
        double n2 = 0.0;
        for (int d = 0; d < DIM; ++d) n2 += c.planar_normal[d] * c.planar_normal[d];
        if (n2 <= 0.0) throw std::runtime_error("planar_regions requires non-zero planar_normal");
    }
    if (c.initial_condition == "explosion" || c.initial_condition == "double_explosion") {
        if (c.explosion_radius <= 0.0) throw std::runtime_error("Invalid explosion radius");
        if (c.initial_condition == "double_explosion" && c.explosion_centers.empty())
            throw std::runtime_error("double_explosion requires explosion_centers");
    }
    if (c.initial_condition == "shock_bubble" || c.initial_condition == "coated_shock_bubble") {
        if (c.shock_axis < 0 || c.shock_axis >= DIM) throw std::runtime_error("Invalid shock axis");
        if (c.bubble_radius <= 0.0) throw std::runtime_error("Invalid bubble radius");
        if (c.left.material < 0 || c.right.material < 0 || c.bubble.material < 0)
            throw std::runtime_error("Invalid material id in shock_bubble IC");
        if (c.initial_condition == "coated_shock_bubble") {
            if (c.film_radius <= c.bubble_radius)
                throw std::runtime_error("coated_shock_bubble requires film_radius > bubble_radius");
            if (c.film.material < 0) throw std::runtime_error("Invalid material id in coated_shock_bubble IC");
        }
    }
}

// load_config (config_loader.hpp:442-570).
template<int DIM>
CaseConfig<DIM> load_case(const std::string& path)
{
    std::ifstream file(path);
    if (!file) throw std::runtime_error("Cannot open config file");
    CaseConfig<DIM> c;
    using Setter = std::function<void(const std::string&)>;
    const auto num = [](double& dst) { return Setter([&dst](const std::string& v) { dst = std::stod(v); }); };
    const auto integer = [](int& dst) { return Setter([&dst](const std::string& v) { dst = std::stoi(v); }); };
    const auto vec = [](std::array<double, DIM>& dst) {
        return Setter([&dst](const std::string& v) { dst = text::doubles<DIM>(v); });
    };
    const auto str = [](std::string& dst) { return Setter([&dst](const std::string& v) { dst = v; }); };
    const auto low = [](std::string& dst) { return Setter([&dst](const std::string& v) { dst = text::lower(v); }); };
    const auto flag = [](bool& dst) { return Setter([&dst](const std::string& v) { dst = text::boolean(v); }); };
    const auto phase = [&](std::unordered_map<std::string, Setter>& t, const std::string& suffix, PhaseState<DIM>& s) {
        t["rho_" + suffix] = num(s.rho);
        t["vel_" + suffix] = vec(s.vel);
        t["p_" + suffix] = num(s.p);
        t["material_" + suffix] = integer(s.material);
    };
    std::unordered_map<std::string, Setter> keys = {
        {"dimension", [](const std::string& v) { if (std::stoi(v) != DIM) throw std::runtime_error("Dimension mismatch"); }},
        {"model_type", [&c](const std::string& v) { model_family(v, false); c.model_type = v; }},
        {"domain_min", vec(c.domain_min)},
        {"domain_max", vec(c.domain_max)},
        {"N", [&c](const std::string& v) { c.N_list.push_back(text::ints<DIM>(v)); }},
        {"tfinal", num(c.tfinal)},
        {"output_times", [&c](const std::string& v) { c.output_times = text::double_list(v); }},
        {"cfl", num(c.cfl)},
        {"exact_riemann", flag(c.exact_riemann)},
        {"output_prefix", str(c.output_prefix)},
        {"output_dir", str(c.output_dir)},
        {"material", [&c](const std::string& v) { c.materials.push_back(parse_material(v)); }},
        {"region", [&c](const std::string& v) { c.regions.push_back(parse_region<DIM>(v)); }},
        {"initial_condition", str(c.initial_condition)},
        {"planar_normal", vec(c.planar_normal)},
        {"explosion_center", vec(c.explosion_center)},
        {"explosion_centers", [&c](const std::string& v) { c.explosion_centers = text::array_list<DIM>(v); }},
        {"explosion_radius", num(c.explosion_radius)},
        {"shock_axis", integer(c.shock_axis)},
        {"shock_position", num(c.shock_position)},
        {"bubble_center", vec(c.bubble_center)},
        {"bubble_radius", num(c.bubble_radius)},
        {"film_radius", num(c.film_radius)},
        {"interface_method", str(c.interface_method)},
        {"time_update", low(c.time_update)},
        {"use_level_set", flag(c.use_level_set)},
        {"reinit_interval", integer(c.reinit_interval)},
        {"reinit_iterations", integer(c.reinit_iterations)},
        {"level_set_reinit_method", low(c.level_set_reinit_method)},
        {"level_set_advection", low(c.level_set_advection)},
        {"level_set_spatial_derivative", low(c.level_set_spatial_derivative)},
        {"level_set_component_policy", low(c.level_set_component_policy)},
        {"rgfm_diagnostics", flag(c.rgfm_diagnostics)},
        {"rgfm_diagnostics_interval", integer(c.rgfm_diagnostics_interval)},
        {"rgfm_star_velocity_mode", low(c.rgfm_star_velocity_mode)},
        {"interface_thickness", num(c.interface_thickness)},
        {"interface_sharpness_alpha", num(c.interface_sharpness_alpha)},
        {"tan_alpha", num(c.interface_sharpness_alpha)},
        {"alpha", num(c.interface_sharpness_alpha)},
        {"dim_alpha_source_floor", num(c.dim_alpha_source_floor)},
        {"dim_lambda_model", low(c.dim_lambda_model)},
        {"barton_solid_material", integer(c.barton_solid_material)},
        {"barton_temperature", num(c.barton_temperature)},
        {"bc_lo", [&c](const std::string& v) { c.bc_lo = text::names<DIM>(v); }},
        {"bc_hi", [&c](const std::string& v) { c.bc_hi = text::names<DIM>(v); }},
    };
    phase(keys, "in", c.in);
    phase(keys, "out", c.out);
    phase(keys, "left", c.left);
    phase(keys, "right", c.right);
    phase(keys, "bubble", c.bubble);
    phase(keys, "film", c.film);

    std::string line;
    while (std::getline(file, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        line = text::trim(line);
        if (line.empty()) continue;
        const auto eq = line.find('=');
        if (eq == std::string::npos) throw std::runtime_error("Invalid line: " + line);
        const std::string key = text::trim(line.substr(0, eq));
        const std::string value = text::trim(line.substr(eq + 1));
        const auto it = keys.find(key);
        if (it == keys.end()) throw std::runtime_error("Unknown key: " + key);
        it->second(value);
    }
    // an automatic model is inferred from interface_method, which rejects unknown methods
    if (model_family(c.model_type, false) == "auto") model_family(c.interface_method, true);
    validate(c);
    return c;
}

// True when the file selects the Barton solid extension (`model = barton_dim`,
// src/dim/barton_dim/driver.hpp:18-32), which this driver does not run.
inline bool selects_barton(const std::string& path)
{
    std::ifstream file(path);
    if (!file) throw std::runtime_error("Cannot open configuration file: " + path);
    std::string line;
    while (std::getline(file, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        line = text::trim(line);
        if (line.empty()) continue;
        const auto eq = line.find('=');
        if (eq == std::string::npos) continue;
        if (text::trim(line.substr(0, eq)) == "model") return text::trim(line.substr(eq + 1)) == "barton_dim";
    }
    return false;
}

}  // namespace dimhost
