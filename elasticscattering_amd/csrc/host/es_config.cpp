// es_config.cpp — .config parsing with the reference's rules
// (src/SimulationConfiguration.h:77-229):
//   * one `key value` per line, split at the first space; lines that are empty
//     or start with '#' or ":/" are skipped; the first occurrence of a key wins;
//   * numbers are read with atoi / atof (leading blanks skipped, trailing text
//     ignored);
//   * defaults: output_type all, output_directory ESLogs (then Result_N),
//     device cpu, target_cell_population 15;
//   * dimension and phi_steps must be odd; magnetic_field_n >= 2; only
//     temperatures > 0 are kept; sample seeds are (seed + 100) * (i + 1).
// Differences: a trailing '\r' is removed from every value, and a missing
// required key is reported instead of throwing std::out_of_range.
#include "es_config.hpp"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <random>
#include <sstream>

namespace es_host {

namespace {

bool need(const std::map<std::string, std::string>& v, const char* key, std::string& out, std::string& err)
{
    auto it = v.find(key);
    if (it == v.end()) {
        err = std::string("missing key '") + key + "'";
        return false;
    }
    out = it->second;
    return true;
}

std::string json_escape(const std::string& s)
{
    std::string o;
    for (char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += c;
        } else if ((unsigned char)c < 0x20) {
            char buf[8];
            std::snprintf(buf, sizeof(buf), "\\u%04x", (unsigned char)c);
            o += buf;
        } else {
            o += c;
        }
    }
    return o;
}

}  // namespace

bool parse_config(const std::string& path, RunConfig& cfg, std::string& error, bool pick_output_dir)
{
    std::ifstream f(path);
    if (!f) {
        error = "Could not load config file '" + path + "'";
        return false;
    }
    std::map<std::string, std::string> values;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '\n' || line[0] == '#' || (line.size() > 1 && line[0] == ':' && line[1] == '/'))
            continue;
        const size_t pos = line.find(' ');
        std::string key = line.substr(0, pos);
        std::string value = (pos == std::string::npos) ? line : line.substr(pos + 1);
        while (!value.empty() && value.back() == '\r') value.pop_back();
        while (!key.empty() && key.back() == '\r') key.pop_back();
        values.emplace(key, value);  // first occurrence wins (unordered_map::insert)
    }

    cfg = RunConfig();
    cfg.config_path = path;
    std::string v;

    cfg.output_type = OutputType::All;
    if (values.count("output_type")) {
        const std::string& lv = values["output_type"];
        if (lv == "nothing") cfg.output_type = OutputType::Nothing;
        else if (lv == "results") cfg.output_type = OutputType::Results;
    }
    if (cfg.output_type != OutputType::Nothing) {
        cfg.base_output_directory = values.count("output_directory") ? values["output_directory"] : "ESLogs";
        if (pick_output_dir) {
            for (unsigned int n = 0;; n++) {
                const std::string dir = cfg.base_output_directory + "/Result_" + std::to_string(n);
                if (!std::filesystem::exists(dir)) {
                    cfg.output_directory = dir;
                    break;
                }
            }
        }
    }
    cfg.use_gpu = values.count("device") ? values["device"] == "gpu" : false;
    cfg.print_info = values.count("print_info") && std::atoi(values["print_info"].c_str()) == 1;

    // Keys this implementation adds.
    if (values.count("cells_per_row")) cfg.cells_per_row = std::atoi(values["cells_per_row"].c_str());
    if (values.count("incoherent_tau")) cfg.incoherent_tau = std::atof(values["incoherent_tau"].c_str());
    if (values.count("impurity_count")) cfg.impurity_count = std::atoll(values["impurity_count"].c_str());
    if (values.count("allow_even")) cfg.allow_even = std::atoi(values["allow_even"].c_str()) == 1;
    if (values.count("magnetic_field_spacing")) cfg.log_fields = values["magnetic_field_spacing"] == "log";
    if (values.count("dump_lifetimes")) cfg.dump_lifetimes = std::atoi(values["dump_lifetimes"].c_str()) == 1;

    if (!need(values, "num_samples", v, error)) return false;
    cfg.num_samples = std::atoi(v.c_str());
    if (cfg.num_samples <= 0) {
        error = "Samples must be greater than 0.";
        return false;
    }
    if (!need(values, "phi_steps", v, error)) return false;
    cfg.particles_per_quadrant = std::atoi(v.c_str());
    if (cfg.particles_per_quadrant % 2 == 0 && !cfg.allow_even) {
        error = "Phi steps must be odd.";
        return false;
    }
    if (!need(values, "dimension", v, error)) return false;
    cfg.positions_per_row = std::atoi(v.c_str());
    if (cfg.positions_per_row % 2 == 0 && !cfg.allow_even) {
        error = "Dimension must be odd.";
        return false;
    }

    if (values.count("magnetic_field")) {
        cfg.magnetic_fields.push_back(std::atof(values["magnetic_field"].c_str()));
    } else {
        std::string a, b, c;
        if (!need(values, "magnetic_field_min", a, error) || !need(values, "magnetic_field_max", b, error) ||
            !need(values, "magnetic_field_n", c, error))
            return false;
        const double mf_min = std::atof(a.c_str()), mf_max = std::atof(b.c_str());
        const double mf_n = std::atoi(c.c_str());
        if (mf_n < 2) {
            error = "Tried to declare magnetic field range with n<2. Use 'magnetic_field x' when using only one value.";
            return false;
        }
        cfg.magnetic_fields.resize((size_t)mf_n);
        if (!cfg.log_fields) {
            const double step = (mf_max - mf_min) / (double)(mf_n - 1);
            for (int i = 0; i < (int)mf_n; i++) cfg.magnetic_fields[i] = mf_min + i * step;
        } else {
            if (!(mf_min > 0) || !(mf_max > 0)) {
                error = "log-spaced magnetic fields need min and max > 0";
                return false;
            }
            const double l0 = std::log10(mf_min), l1 = std::log10(mf_max);
            for (int i = 0; i < (int)mf_n; i++) cfg.magnetic_fields[i] = std::pow(10.0, l0 + (l1 - l0) * i / (mf_n - 1));
        }
    }

    if (!need(values, "temperatures", v, error)) return false;
    {
        std::stringstream ss(v);
        while (ss.good()) {
            std::string a;
            std::getline(ss, a, ' ');
            const double temp = std::atof(a.c_str());
            if (temp > 0 || (cfg.incoherent_tau > 0 && temp == 0 && !a.empty() && a.find_first_of("0123456789") != std::string::npos))
                cfg.temperatures.push_back(temp);
        }
    }

    if (!need(values, "particle_speed", v, error)) return false;
    cfg.particle_speed = std::atof(v.c_str());
    if (!need(values, "tau", v, error)) return false;
    cfg.tau = std::atof(v.c_str());
    if (!need(values, "alpha", v, error)) return false;
    cfg.alpha = std::atof(v.c_str());
    if (!need(values, "clockwise", v, error)) return false;
    cfg.is_clockwise = std::atoi(v.c_str()) > 0;

    if (values.count("start_seed")) {
        cfg.start_seed = (unsigned int)std::atoi(values["start_seed"].c_str());
        cfg.seed_given = true;
    } else {
        std::random_device rd;
        cfg.start_seed = rd();
    }
    cfg.sample_seeds.resize(cfg.num_samples);
    for (int i = 0; i < cfg.num_samples; i++) cfg.sample_seeds[i] = (cfg.start_seed + 100u) * (unsigned int)(i + 1);

    if (!need(values, "region_size", v, error)) return false;
    cfg.region_size = std::atof(v.c_str());
    if (!need(values, "region_extends", v, error)) return false;
    cfg.region_extends = std::atof(v.c_str());
    if (!need(values, "impurity_density", v, error) && cfg.impurity_count <= 0) return false;
    cfg.impurity_density = values.count("impurity_density") ? std::atof(values["impurity_density"].c_str()) : 0.0;
    error.clear();
    if (!need(values, "impurity_radius", v, error)) return false;
    cfg.impurity_radius = std::atof(v.c_str());
    cfg.target_cell_population =
        values.count("target_cell_population") ? std::atoi(values["target_cell_population"].c_str()) : 15;

    int x = cfg.num_samples;
    cfg.digits_in_sample_num = 0;
    while (x > 0) {
        x = x / 10;
        cfg.digits_in_sample_num++;
    }
    return true;
}

std::string config_to_json(const RunConfig& c)
{
    std::ostringstream o;
    o.precision(17);
    auto vec = [&](const std::vector<double>& v) {
        o << "[";
        for (size_t i = 0; i < v.size(); i++) o << (i ? ", " : "") << v[i];
        o << "]";
    };
    o << "{\"num_samples\": " << c.num_samples << ", \"positions_per_row\": " << c.positions_per_row
      << ", \"particles_per_quadrant\": " << c.particles_per_quadrant << ", \"print_info\": " << (c.print_info ? 1 : 0)
      << ", \"output_type\": \""
      << (c.output_type == OutputType::Nothing ? "nothing" : c.output_type == OutputType::Results ? "results" : "all")
      << "\", \"use_gpu\": " << (c.use_gpu ? 1 : 0) << ", \"tau\": " << c.tau << ", \"alpha\": " << c.alpha
      << ", \"particle_speed\": " << c.particle_speed << ", \"is_clockwise\": " << (c.is_clockwise ? 1 : 0)
      << ", \"region_size\": " << c.region_size << ", \"region_extends\": " << c.region_extends
      << ", \"impurity_density\": " << c.impurity_density << ", \"impurity_radius\": " << c.impurity_radius
      << ", \"target_cell_population\": " << c.target_cell_population << ", \"magnetic_fields\": ";
    vec(c.magnetic_fields);
    o << ", \"temperatures\": ";
    vec(c.temperatures);
    o << ", \"sample_seeds\": [";
    for (size_t i = 0; i < c.sample_seeds.size(); i++) o << (i ? ", " : "") << c.sample_seeds[i];
    o << "], \"start_seed\": " << c.start_seed << ", \"seed_given\": " << (c.seed_given ? 1 : 0)
      << ", \"output_directory\": \"" << json_escape(c.output_directory) << "\", \"base_output_directory\": \""
      << json_escape(c.base_output_directory) << "\", \"digits_in_sample_num\": " << c.digits_in_sample_num
      << ", \"cells_per_row\": " << c.cells_per_row << ", \"incoherent_tau\": " << c.incoherent_tau
      << ", \"impurity_count\": " << c.impurity_count << ", \"allow_even\": " << (c.allow_even ? 1 : 0)
      << ", \"log_fields\": " << (c.log_fields ? 1 : 0) << ", \"dump_lifetimes\": " << (c.dump_lifetimes ? 1 : 0)
      << "}";
    return o.str();
}

}  // namespace es_host
