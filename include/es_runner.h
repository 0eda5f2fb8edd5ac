/*
 * es_runner.h — C entry points of the host runner (the part of the
 * reference's CLI that sits above the Simulation backend):
 *   .config parsing   src/SimulationConfiguration.h:77-229
 *   the run           src/SimulationRunner.cpp:32-198
 *   result files      src/Logger.cpp:42-272 (T{j}.dat, sample_results.dat, Metrics.txt)
 * The runner calls the solver through include/es_abi.h on one GPU.
 */
#ifndef ES_RUNNER_H
#define ES_RUNNER_H

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Parses a .config file and writes the resulting configuration as one JSON
 * object into out (cap bytes).  Returns ES_OK, ES_ERR_INVALID_ARG with the
 * reason in the JSON's "error" field, or a negative size when out is too small.
 */
int es_config_parse(const char* path, char* out, int cap);

/*
 * SimulationRunner::Run (SimulationRunner.cpp:32-52): parse, create the output
 * directory, run every sample (coherent then incoherent pass, each field), and
 * write the result files.  verbose=0 prints nothing.  Writes the output
 * directory used into out_dir (cap bytes) when out_dir is not NULL.
 */
int es_run(const char* config_path, int verbose, char* out_dir, int cap);

/*
 * Logger::LogImages (Logger.cpp:320-334): writes the three dim x dim images of
 * one temperature (phi-integrated lifetime, sigma_xx, sigma_xy; as returned by
 * es_solver_compute_sigmas_with_images) side by side as an 8-bit RGB PNG.
 */
int es_log_images(const char* path, int dim, const double* lifetimes, const double* sigma_xx,
                  const double* sigma_xy);

#ifdef __cplusplus
}
#endif

#endif /* ES_RUNNER_H */
