// es_runner.cpp — the host runner above the solver: the behaviour of the
// reference's SimulationRunner (src/SimulationRunner.cpp:32-221) and the
// result files of its Logger (src/Logger.cpp:42-272), driving the GPU solver
// through include/es_abi.h.
//
// Result files keep the reference's layout (see the format notes at each
// writer).  Metrics.txt is written as UTF-8 with the "C" number format: the
// reference imbues en_US.UTF-8 (Logger.cpp:142), which groups digits of
// numbers >= 1000 with ',' and throws where that locale is missing.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/es_abi.h"
#include "../../../include/es_runner.h"
#include "es_config.hpp"
#include "es_png.hpp"

namespace es_host {

namespace {

constexpr int kResultPrecision = 6;        // Logger.h:31
constexpr int kRw = kResultPrecision + 10;  // Logger.h:32

struct SigmaPair {
    double xx = 0, xy = 0;
};

// SampleResult (SimulationResult.h:51-62): [field][temperature].
using SampleResult = std::vector<std::vector<SigmaPair>>;

// SampleMetrics (Metrics.h:55-80).
struct SampleMetrics {
    int sample_index = 0;
    long long total_particles = 0;
    long long total_cells = 0;
    long long impurity_count = 0;
    unsigned int seed = 0;
    long long total_indexed_impurities = 0;
    bool coherent = false;
    std::vector<es_metrics> iterations;
};

std::string result_path(const RunConfig& c, int t) { return c.output_directory + "/T" + std::to_string(t) + ".dat"; }
std::string sample_path(const RunConfig& c, int s) { return c.output_directory + "/Sample " + std::to_string(s); }
std::string metrics_path(const RunConfig& c) { return c.output_directory + "/Metrics.txt"; }

// Logger::CreateResultLog (Logger.cpp:42-66).
void write_result_header(const std::string& path, const RunConfig& c, double temperature)
{
    std::ofstream f(path);
    f << "# Elastic Scattering results summary.\n\n";
    f << "Temperature " << std::fixed << std::setprecision(4) << temperature << "\n\n";
    f << "MF values   " << c.magnetic_fields.size() << "\n";
    f << "Samples     " << c.num_samples << "\n";
    f << "Positions   " << c.positions_per_row << " x " << c.positions_per_row << "\n";
    f << "Phi steps   " << c.particles_per_quadrant << "\n\n";
    f << "# Results #\n\n";
    f << std::left << std::setw(8) << "B" << std::setw(kRw) << "sigma_xx_inc" << std::setw(kRw) << "sigma_xx_coh"
      << std::setw(kRw) << "sigma_xy_inc" << std::setw(kRw) << "sigma_xy_coh" << std::setw(kRw) << "delta_xx" << "\n";
}

// Logger::LogResult (Logger.cpp:68-80).
void append_result_row(const std::string& path, double B, const SigmaPair& inc, const SigmaPair& coh, double xxd)
{
    std::ofstream f(path, std::ios_base::app);
    f << std::left << std::setw(8) << std::setprecision(2) << std::fixed << B << std::scientific
      << std::setprecision(kResultPrecision) << std::setw(kRw) << inc.xx << std::setw(kRw) << coh.xx << std::setw(kRw)
      << inc.xy << std::setw(kRw) << coh.xy << std::setw(kRw) << std::setprecision(4) << std::defaultfloat << xxd
      << "\n";
}

// Logger::CreateSampleResultLog / LogSampleResults (Logger.cpp:82-121).
void write_sample_results(const std::string& path, const SampleResult& coh, const SampleResult& inc)
{
    std::ofstream f(path);
    f << "# Elastic Scattering sample results summary.\n\n";
    f << std::left << std::setw(4) << "T" << std::setw(4) << "B" << std::setw(kRw) << "sigma_xx_inc" << std::setw(kRw)
      << "sigma_xx_coh" << std::setw(kRw) << "sigma_xy_inc" << std::setw(kRw) << "sigma_xy_coh" << "\n";
    f << std::setprecision(kResultPrecision);
    const size_t nB = coh.size(), nT = nB ? coh[0].size() : 0;
    for (size_t j = 0; j < nT; j++) {
        for (size_t i = 0; i < nB; i++) {
            f << std::left << std::setw(4) << j << std::setw(4) << i << std::setw(kRw) << std::scientific << inc[i][j].xx
              << std::setw(kRw) << std::scientific << coh[i][j].xx << std::setw(kRw) << std::scientific << inc[i][j].xy
              << std::setw(kRw) << std::scientific << coh[i][j].xy << "\n";
        }
        f << "\n";
    }
}

// Logger::CreateSampleMetricsLog (Logger.cpp:123-136).
void write_metrics_header(const std::string& path, const RunConfig& c, long long unique, int cpr, double grid_seconds)
{
    std::ofstream f(path);
    f << "# Metrics collected during particle lifetime computation.\n\n";
    f << "Positions  " << c.positions_per_row << " x " << c.positions_per_row << "\n";
    f << "Phi values " << 4 * c.particles_per_quadrant << "\n";
    f << "Lifetimes  " << (long long)c.positions_per_row * c.positions_per_row * 4 * c.particles_per_quadrant << "\n";
    f << "Impurities " << unique << "\n";
    f << "Cells      " << cpr << " x " << cpr << "\n";
    f << "Index time " << std::fixed << std::setprecision(1) << grid_seconds * 1000.0 << "ms\n\n";
}

std::string rep(const char* utf8, int n)
{
    std::string s;
    for (int i = 0; i < n; i++) s += utf8;
    return s;
}

// Logger::LogSampleMetrics (Logger.cpp:138-272): the per-field table.
void append_sample_metrics(const std::string& path, const SampleMetrics& sm)
{
    const char* H = "─";
    const int mw = 26, vw = 12;
    const auto& m = sm.iterations;
    std::ostringstream f;
    if (sm.coherent) {
        const double inside = m.empty() ? 0.0 : (double)m[0].particles_inside_impurity;
        f << "\n\n\n";
        f << "---------------------------------------------------------\n";
        f << "Sample " << sm.sample_index << (sm.coherent ? " Coherent" : " Incoherent") << "\n";
        f << "---------------------------------------------------------\n";
        f << "Seed                      " << sm.seed << "\n";
        f << "Total indexed impurities  " << std::setprecision(2) << (double)sm.total_indexed_impurities / 1000000
          << " mil (" << std::setprecision(4) << (sm.total_indexed_impurities - sm.impurity_count) << " overlapped)\n";
        f << "Avg. impurities per cell  " << std::setprecision(2)
          << (double)sm.total_indexed_impurities / (double)sm.total_cells << "\n";
        f << "Positions inside impurity " << std::setprecision(2) << (100 * (inside / (double)sm.total_particles))
          << "%\n";
    }
    const size_t n = m.size();
    f << "\n┌" << rep(H, mw);
    for (size_t i = 0; i < n; i++) f << "┬" << rep(H, vw);
    f << "┐\n";
    f << "│" << std::setw(mw) << std::left << (sm.coherent ? " Metric (Coherent)" : " Metric (Incoherent)");
    for (size_t i = 0; i < n; i++) f << "│" << std::setw(vw) << std::left << (" MF " + std::to_string(i));
    f << "│\n";
    f << "├" << rep(H, mw);
    for (size_t i = 0; i < n; i++) f << "┼" << rep(H, vw);
    f << "┤\n";

    auto row = [&](const char* label, int lw, auto value_fn) {
        f << "│" << std::setw(lw) << std::left << label;
        for (size_t i = 0; i < n; i++) value_fn(i);
        f << "│\n";
    };
    auto blank = [&]() {
        f << "│" << std::string(mw, ' ');
        for (size_t i = 0; i < n; i++) f << "│" << std::string(vw, ' ');
        f << "│\n";
    };
    row(" Avg. particle lifetime", mw, [&](size_t i) {
        f << "│" << std::setw(vw) << std::right << std::setprecision(3) << m[i].avg_particle_lifetime;
    });
    row("Time spent on", mw, [&](size_t) { f << "│" << std::string(vw, ' '); });
    row("\t Lifetimes", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw - 2) << std::right << std::setprecision(3) << (m[i].time_elapsed_lifetimes * 1000.0)
          << "ms";
    });
    row("\t Temperatures", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw - 2) << std::right << std::setprecision(3)
          << (m[i].time_elapsed_temperatures * 1000.0) << "ms";
    });
    row(" Particles escaped", mw, [&](size_t i) {
        f << "│" << std::setw(vw) << std::right << (double)m[i].particles_escaped;
    });
    row(" Particles at bound", mw, [&](size_t i) {
        f << "│" << std::setw(vw) << std::right << (double)m[i].particles_at_bound;
    });
    blank();
    row(" Total impurities tested", mw, [&](size_t i) {
        f << "│" << std::setw(vw - 4) << std::right << std::setprecision(3)
          << ((double)m[i].impurities_tested / 1000000.0) << " mil";
    });
    row("\t Average", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw) << std::right << std::setprecision(2)
          << ((double)m[i].impurities_tested / (double)m[i].real_particles);
    });
    row("\t Average %", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw - 1) << std::right << std::setprecision(2)
          << (100 * ((double)m[i].impurities_tested / (double)m[i].real_particles) / (double)sm.impurity_count) << "%";
    });
    blank();
    row(" Grid cells passed", mw, [&](size_t i) {
        f << "│" << std::setw(vw - 4) << std::right << std::setprecision(3)
          << ((double)m[i].cells_passed / 1000000.0) << " mil";
    });
    row("\t Average", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw) << std::right << std::setprecision(2)
          << ((double)m[i].cells_passed / (double)m[i].real_particles);
    });
    row("\t Average %", mw - 2, [&](size_t i) {
        f << "│" << std::setw(vw - 1) << std::right << std::setprecision(2)
          << (100 * ((double)m[i].cells_passed / (double)m[i].real_particles) / (double)sm.total_cells) << "%";
    });
    f << "└" << rep(H, mw);
    for (size_t i = 0; i < n; i++) f << "┴" << rep(H, vw);
    f << "┘\n";
    std::ofstream out(path, std::ios_base::app);
    out << f.str();
}

struct Runner {
    RunConfig cfg;
    es_solver* s = nullptr;
    bool verbose = true;
    std::string err;

    ~Runner()
    {
        if (s) es_solver_destroy(s);
    }

    int check(int rc)
    {
        if (rc != ES_OK) err = es_last_error(s);
        return rc;
    }

    es_grid_desc grid_desc() const
    {
        es_grid_desc d{};
        d.region_size = cfg.region_size;
        d.region_extends = cfg.region_extends;
        d.impurity_radius = cfg.impurity_radius;
        d.impurity_density = cfg.impurity_density;
        d.target_cell_population = cfg.target_cell_population;
        d.cells_per_row = cfg.cells_per_row;
        return d;
    }

    // SimulationRunner::RunSample (SimulationRunner.cpp:103-151).
    int run_sample(int sample_index, bool coherent, const es_grid_info& gi, SampleResult& out)
    {
        const int nT = (int)cfg.temperatures.size(), nB = (int)cfg.magnetic_fields.size();
        es_settings st{};
        st.tau = cfg.tau;
        st.alpha = cfg.alpha;
        st.is_clockwise = cfg.is_clockwise ? 1 : 0;
        st.particle_speed = cfg.particle_speed;
        st.incoherent_tau = cfg.incoherent_tau;
        int rc = check(es_solver_init_sample(s, &st, coherent ? 1 : 0));
        if (rc) return rc;
        SampleMetrics sm;
        sm.sample_index = sample_index;
        sm.coherent = coherent;
        sm.total_particles = (long long)cfg.positions_per_row * cfg.positions_per_row * 4 * cfg.particles_per_quadrant;
        sm.total_indexed_impurities = gi.indexed_impurity_count;
        sm.impurity_count = gi.unique_impurity_count;
        sm.total_cells = (long long)gi.cells_per_row * gi.cells_per_row;
        sm.seed = cfg.sample_seeds[sample_index];
        sm.iterations.resize(nB);
        if (verbose) {
            const std::string raw = std::to_string(sample_index + 1);
            std::cout << "\r" << std::string(cfg.digits_in_sample_num - raw.size(), '0') << raw
                      << (coherent ? " C" : " I") << " [" << std::flush;
        }
        out.assign(nB, std::vector<SigmaPair>(nT));
        std::vector<es_sigma> sig((size_t)nB * std::max(nT, 1));
        if (cfg.output_type == OutputType::All && !cfg.dump_lifetimes) {
            // ComputeSigmasWithImages + LogImages per field (SimulationRunner.cpp:134-141)
            const size_t plane = (size_t)cfg.positions_per_row * cfg.positions_per_row;
            std::vector<double> img((size_t)nT * 3 * plane);
            for (int b = 0; b < nB; b++) {
                rc = check(es_solver_compute_sigmas_with_images(s, cfg.magnetic_fields[b], cfg.temperatures.data(), nT,
                                                                sig.data() + (size_t)b * nT, &sm.iterations[b],
                                                                img.data()));
                if (rc) return rc;
                for (int t = 0; t < nT; t++) {
                    const double* im = img.data() + (size_t)t * 3 * plane;
                    const std::string path = sample_path(cfg, sample_index) + (coherent ? "/Coherent/" : "/Incoherent/") +
                                             "T" + std::to_string(t) + " MF" + std::to_string(b) + ".png";
                    if (!es_host::log_images(path, cfg.positions_per_row, im, im + plane, im + 2 * plane)) {
                        err = "cannot write " + path;
                        return ES_ERR_INVALID_ARG;
                    }
                }
            }
        } else if (!cfg.dump_lifetimes) {
            rc = check(es_solver_compute_sigmas_sweep(s, cfg.magnetic_fields.data(), nB, cfg.temperatures.data(), nT,
                                                      sig.data(), sm.iterations.data()));
            if (rc) return rc;
        } else {
            const size_t P = (size_t)sm.total_particles;
            std::vector<double> lts(P);
            const std::string dir = sample_path(cfg, sample_index) + "/Lifetimes";
            std::filesystem::create_directories(dir);
            for (int b = 0; b < nB; b++) {
                rc = check(es_solver_compute_sigmas(s, cfg.magnetic_fields[b], cfg.temperatures.data(), nT,
                                                    sig.data() + (size_t)b * nT, &sm.iterations[b], lts.data()));
                if (rc) return rc;
                std::ofstream lf(dir + (coherent ? "/C MF" : "/I MF") + std::to_string(b) + ".f64", std::ios::binary);
                lf.write(reinterpret_cast<const char*>(lts.data()), (std::streamsize)(P * sizeof(double)));
            }
        }
        for (int b = 0; b < nB; b++)
            for (int t = 0; t < nT; t++) {
                out[b][t].xx = sig[(size_t)b * nT + t].xx;
                out[b][t].xy = sig[(size_t)b * nT + t].xy;
            }
        if (verbose) std::cout << std::string(nB, 'x') << "]" << std::flush;
        if (cfg.output_type != OutputType::Nothing) append_sample_metrics(metrics_path(cfg), sm);
        return ES_OK;
    }

    // SimulationRunner::Run / RunSimulation / FinishResults (SimulationRunner.cpp:32-198).
    int run()
    {
        namespace fs = std::filesystem;
        if (cfg.temperatures.empty()) {
            err = "no temperature > 0 in the configuration";
            return ES_ERR_INVALID_ARG;
        }
        if (verbose) std::cout << "Starting simulation (" << cfg.config_path << ")" << std::endl;
        if (cfg.output_type != OutputType::Nothing) {
            if (verbose) std::cout << "Output directory: " << cfg.output_directory << std::endl;
            std::error_code ec;
            fs::create_directories(cfg.base_output_directory, ec);
            fs::create_directory(cfg.output_directory, ec);
            const std::string name = fs::path(cfg.config_path).filename().string();
            fs::copy_file(cfg.config_path, cfg.output_directory + "/" + name, fs::copy_options::overwrite_existing, ec);
        }
        const auto t_begin = std::chrono::steady_clock::now();

        es_solver_desc desc{};
        desc.positions_per_row = cfg.positions_per_row;
        desc.particles_per_quadrant = cfg.particles_per_quadrant;
        desc.precision = ES_PRECISION_FP64;
        desc.device = 0;
        int rc = es_solver_create(&desc, &s);
        if (rc != ES_OK) {
            err = es_last_error(nullptr);
            return rc;
        }
        const es_grid_desc gd = grid_desc();
        es_grid_info gi{};
        const auto tg0 = std::chrono::steady_clock::now();
        rc = check(es_grid_generate(s, cfg.sample_seeds[0], cfg.impurity_count, &gd, &gi));
        if (rc) return rc;
        const double grid_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tg0).count();
        if (cfg.output_type != OutputType::Nothing)
            write_metrics_header(metrics_path(cfg), cfg, gi.unique_impurity_count, gi.cells_per_row, grid_seconds);

        std::vector<SampleResult> coh(cfg.num_samples), inc(cfg.num_samples);
        for (int i = 0; i < cfg.num_samples; i++) {
            if (i > 0) {
                rc = check(es_grid_generate(s, cfg.sample_seeds[i], cfg.impurity_count, &gd, &gi));
                if (rc) return rc;
            }
            if (cfg.output_type == OutputType::All) {
                const std::string sp = sample_path(cfg, i);
                fs::create_directories(sp + "/Incoherent/");
                fs::create_directories(sp + "/Coherent/");
            }
            if ((rc = run_sample(i, true, gi, coh[i]))) return rc;
            if ((rc = run_sample(i, false, gi, inc[i]))) return rc;
            if (cfg.output_type == OutputType::All)
                write_sample_results(sample_path(cfg, i) + "/sample_results.dat", coh[i], inc[i]);
        }
        const double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_begin).count();
        if (verbose) std::cout << "\n\nFinished! (" << total << "s)" << std::endl;
        if (cfg.output_type != OutputType::Nothing) finish(coh, inc);
        return ES_OK;
    }

    // SimulationRunner::FinishResults (SimulationRunner.cpp:153-198).
    void finish(const std::vector<SampleResult>& coh_s, const std::vector<SampleResult>& inc_s)
    {
        const double S = (double)cfg.num_samples;
        for (size_t j = 0; j < cfg.temperatures.size(); j++) {
            const std::string path = result_path(cfg, (int)j);
            write_result_header(path, cfg, cfg.temperatures[j]);
            for (size_t i = 0; i < cfg.magnetic_fields.size(); i++) {
                SigmaPair coherent, incoherent;
                double dxx_squared = 0;
                for (int s_ = 0; s_ < cfg.num_samples; s_++) {
                    const SigmaPair c = coh_s[s_][i][j], n = inc_s[s_][i][j];
                    coherent.xx += c.xx;
                    coherent.xy += c.xy;
                    incoherent.xx += n.xx;
                    incoherent.xy += n.xy;
                    const double v = c.xx + n.xx;
                    dxx_squared += v * v;
                }
                coherent.xx /= S;
                coherent.xy /= S;
                incoherent.xx /= S;
                incoherent.xy /= S;
                double xxd = 0;
                if (cfg.num_samples != 1) {
                    const double sxx_sq_exp = dxx_squared / S;
                    const double sxx_exp = coherent.xx + incoherent.xx;
                    const double sxx_std = std::sqrt((sxx_sq_exp - sxx_exp * sxx_exp) / (S - 1.0));
                    xxd = sxx_std / sxx_exp;
                }
                append_result_row(path, cfg.magnetic_fields[i], incoherent, coherent, xxd);
            }
        }
    }
};

}  // namespace

}  // namespace es_host

extern "C" int es_log_images(const char* path, int dim, const double* lifetimes, const double* sigma_xx,
                             const double* sigma_xy)
{
    if (!path || dim <= 0 || !lifetimes || !sigma_xx || !sigma_xy) return ES_ERR_INVALID_ARG;
    return es_host::log_images(path, dim, lifetimes, sigma_xx, sigma_xy) ? ES_OK : ES_ERR_INVALID_ARG;
}

extern "C" int es_config_parse(const char* path, char* out, int cap)
{
    es_host::RunConfig cfg;
    std::string err;
    const bool ok = path && es_host::parse_config(path, cfg, err, false);
    std::string json;
    if (ok) {
        json = es_host::config_to_json(cfg);
    } else {
        json = "{\"error\": \"";
        for (char c : err) json += (c == '"' ? '\'' : c);
        json += "\"}";
    }
    if (!out || cap < (int)json.size() + 1) return -(int)(json.size() + 1);
    std::memcpy(out, json.c_str(), json.size() + 1);
    return ok ? ES_OK : ES_ERR_INVALID_ARG;
}

extern "C" int es_run(const char* config_path, int verbose, char* out_dir, int cap)
{
    es_host::Runner r;
    r.verbose = verbose != 0;
    std::string err;
    if (!config_path || !es_host::parse_config(config_path, r.cfg, err, true)) {
        if (verbose) std::cerr << err << std::endl;
        return ES_ERR_INVALID_ARG;
    }
    if (out_dir && cap > 0) {
        std::snprintf(out_dir, (size_t)cap, "%s", r.cfg.output_directory.c_str());
    }
    if (verbose && !r.cfg.use_gpu)
        std::cout << "note: device cpu requested; this implementation runs the solver on the GPU" << std::endl;
    const int rc = r.run();
    if (rc != ES_OK && verbose) std::cerr << "error: " << r.err << std::endl;
    return rc;
}
