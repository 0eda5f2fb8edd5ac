// es_config.hpp — the .config format of the reference (SimulationConfiguration.h).
#pragma once

#include <string>
#include <vector>

namespace es_host {

enum class OutputType { Nothing, Results, All };

// SimulationConfiguration (SimulationConfiguration.h:54-76) + Settings
// (ConfigSettings.h:22-34), plus this implementation's optional keys.
struct RunConfig {
    int num_samples = 0;
    int positions_per_row = 0;
    int particles_per_quadrant = 0;
    bool print_info = false;
    OutputType output_type = OutputType::All;
    bool use_gpu = false;            // the `device` key; the solver always runs on the GPU

    double tau = 0, alpha = 0, particle_speed = 0;
    bool is_clockwise = false;
    double region_size = 0, region_extends = 0, impurity_density = 0, impurity_radius = 0;
    int target_cell_population = 15;

    std::vector<double> magnetic_fields;
    std::vector<double> temperatures;
    std::vector<unsigned int> sample_seeds;
    unsigned int start_seed = 0;
    bool seed_given = false;

    std::string config_path;
    std::string base_output_directory;
    std::string output_directory;
    int digits_in_sample_num = 0;

    // Keys this implementation adds (absent: reference behaviour).
    int cells_per_row = 0;           // explicit cell grid
    double incoherent_tau = 0;       // tau of the incoherent pass; temperatures may then be 0
    long long impurity_count = 0;    // explicit count instead of ceil(side^2 * density)
    bool allow_even = false;         // accept even dimension / phi_steps
    bool log_fields = false;         // magnetic_field_spacing log
    bool dump_lifetimes = false;     // write raw lifetimes per sample, pass and field
};

// Parses `path`; on failure returns false with the reference's message (or the
// missing key) in `error`.  `pick_output_dir` chooses <base>/Result_N.
bool parse_config(const std::string& path, RunConfig& cfg, std::string& error, bool pick_output_dir = true);

std::string config_to_json(const RunConfig& cfg);

}  // namespace es_host
