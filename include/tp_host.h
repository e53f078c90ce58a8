/*
 * tp_host.h -- host side of the 3photons event loop around the MI355X engine:
 * configuration, constants, finalisation and output.  These are the parts the
 * reference keeps on the CPU (src/config.rs, src/coupling.rs, src/resacc.rs
 * new/finalize, src/resfin.rs, src/output.rs); they are O(1) per run and are
 * restated here in C++ behind a C ABI so that a caller -- the trois_photons
 * CLI, a Rust crate, or Python via ctypes -- gets the reference's numbers and
 * files from the engine's batch accumulators.
 */
#ifndef TP_HOST_H
#define TP_HOST_H

#include <stdint.h>

#include "tp.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The 18 positional items of `valeurs` (config.rs:8-53). */
typedef struct {
    uint64_t num_events;
    double e_total, cut_beam, cut_pp, e_min, cut_plane;
    double alpha, alpha_z, gev2_to_picobarn, m_z0, g_z0, sin2_weinberg, branching_ep_em;
    double beta_plus, beta_minus;
    int32_t num_bins, impr, plot;
    int32_t parsed; /* 1 once all 18 items parsed, even if a validity check then fails */
} tp_config;

/* Final results (resfin.rs:26-62), [spin][contribution]. */
typedef struct {
    uint64_t selected_events;
    double spm2[2][5], vars[2][5];
    double sigma, prec, variance, beta_min, ss_p, inc_ss_p, ss_m, inc_ss_m;
} tp_final;

/* Configuration::load (config.rs:57-128): first token of each non-blank line,
 * Fortran booleans, then the three ensure! checks.  Returns TP_OK or
 * TP_ERR_CONFIG with a message in err (err_len bytes). */
int tp_config_load(const char *path, tp_config *cfg, char *err, int err_len);

/* Configuration's Display (config.rs:131-155) into buf; returns bytes needed. */
int tp_config_format(const tp_config *cfg, char *buf, int buf_len);

/* Couplings::new + EventGenerator::new + ResultsAccumulator::new constants
 * (coupling.rs:23-33, evgen.rs:40-77, resacc.rs:59-100) into *params, with
 * sort_photons = 1 and precision = TP_PREC_F64. */
int tp_params_from_config(const tp_config *cfg, tp_params *params);

/* ResultsAccumulator::merge (resacc.rs:133-139): into += other. */
void tp_acc_merge(tp_acc *into, const tp_acc *other);

/* ResultsAccumulator::finalize (resacc.rs:142-223). */
int tp_finalize(const tp_config *cfg, const tp_acc *merged, tp_final *out);

/* The same three with the reference's `f32` feature (numeric.rs:6-14): the
 * configuration values rounded to f32 and every constant, sum and final
 * quantity computed in f32 arithmetic (results stored as doubles holding f32
 * values); tp_params_from_config_f32 sets precision = TP_PREC_F32. */
int tp_params_from_config_f32(const tp_config *cfg, tp_params *params);
void tp_acc_merge_f32(tp_acc *into, const tp_acc *other);
int tp_finalize_f32(const tp_config *cfg, const tp_acc *merged, tp_final *out);

/* Outputs of an f32 run: the values printed as f32 (Display as the shortest
 * f32 round trip, write_engineering with f32::DIGITS - 1 = 5 significant
 * digits, output.rs:26,180-269) and the report computed in f32. */
int tp_config_format_f32(const tp_config *cfg, char *buf, int buf_len);
int tp_format_res_data_f32(const tp_config *cfg, const tp_final *res, char *buf, int buf_len);
int tp_format_report_f32(const tp_config *cfg, const tp_final *res, char *buf, int buf_len);
int tp_dump_results_f32(const tp_config *cfg, const tp_final *res, double elapsed_s, const char *dir);

/* Text of res.data (output.rs:64-142) into buf; returns bytes needed. */
int tp_format_res_data(const tp_config *cfg, const tp_final *res, char *buf, int buf_len);

/* Text of FinalResults::eric + fawzi (resfin.rs:66-194) into buf. */
int tp_format_report(const tp_config *cfg, const tp_final *res, char *buf, int buf_len);

/* output::dump_results (output.rs:30-177): prints the report on stdout and
 * writes res.times, res.data and appends pil.mc in directory dir (NULL = cwd).
 * elapsed_s is the run's wall time (main.rs:85,138). */
int tp_dump_results(const tp_config *cfg, const tp_final *res, double elapsed_s, const char *dir);

#ifdef __cplusplus
}
#endif
#endif /* TP_HOST_H */
