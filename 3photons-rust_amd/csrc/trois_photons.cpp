// trois_photons.cpp -- command-line driver mirroring the reference's main()
// (HadrienG2/3photons-rust src/main.rs:75-145) with the event loop on one
// MI355X through the C ABI of include/tp.h.
//
// The reference selects its variants with cargo features; here they are flags:
//   --standard-random     xoshiro256+ stream instead of RANF
//   --faster-threading    per-batch jump()/reseed instead of skip-ahead
//   --no-photon-sorting   keep RAMBO's photon order
//   --faster-evgen        the reference's faster-evgen generation (any stream, f64 or f32)
//   --mixed               f32 spinor/amplitude arithmetic (tp_precision MIXED)
//   --f32                 the reference's f32 feature: everything in f32,
//                         xoshiro128+ with --standard-random, f32 outputs
//   --config PATH         configuration file (default: ./valeurs)
//   --device N            HIP device ordinal (default 0)
// Outputs are the reference's: stdout report, res.data, res.times, pil.mc in
// the working directory.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "tp.h"
#include "tp_host.h"

int main(int argc, char **argv)
{
    std::string config = "valeurs";
    int device = 0, rng = TP_RNG_RANF, seeding = TP_SEED_SKIPAHEAD, sort = 1, precision = TP_PREC_F64;
    int faster_evgen = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--standard-random")
            rng = TP_RNG_XOSHIRO;
        else if (a == "--faster-threading")
            seeding = TP_SEED_JUMP;
        else if (a == "--no-photon-sorting")
            sort = 0;
        else if (a == "--faster-evgen")
            faster_evgen = 1;
        else if (a == "--mixed")
            precision = TP_PREC_MIXED;
        else if (a == "--f32")
            precision = TP_PREC_F32;
        else if (a == "--config" && i + 1 < argc)
            config = argv[++i];
        else if (a == "--device" && i + 1 < argc)
            device = std::atoi(argv[++i]);
        else {
            std::fprintf(stderr, "usage: %s [--standard-random] [--faster-threading] [--no-photon-sorting] "
                                 "[--faster-evgen] [--mixed] [--f32] [--config PATH] [--device N]\n",
                         argv[0]);
            return 2;
        }
    }
    // ### CONFIGURATION READOUT ### (main.rs:77-79, config.rs:57-128)
    tp_config cfg;
    char err[512];
    int rc = tp_config_load(config.c_str(), &cfg, err, sizeof err);
    if (cfg.parsed) { // the reference echoes the configuration before its ensure! checks
        auto format = precision == TP_PREC_F32 ? tp_config_format_f32 : tp_config_format;
        std::vector<char> text(static_cast<size_t>(format(&cfg, nullptr, 0)));
        format(&cfg, text.data(), static_cast<int>(text.size()));
        std::fputs(text.data(), stdout);
    }
    if (rc != TP_OK) {
        std::fprintf(stderr, "Error: failed to load the configuration\n\nCaused by:\n    %s\n", err);
        return 1;
    }
    // ### SIMULATION INITIALIZATION ### (main.rs:85-97)
    auto t0 = std::chrono::steady_clock::now();
    tp_params params;
    if (precision == TP_PREC_F32) {
        tp_params_from_config_f32(&cfg, &params);
        if (rng == TP_RNG_XOSHIRO)
            rng = TP_RNG_XOSHIRO128; // the f32 feature's generator (standard.rs:8-11)
    } else {
        tp_params_from_config(&cfg, &params);
    }
    params.sort_photons = sort;
    params.precision = precision;
    params.faster_evgen = faster_evgen;
    std::puts("IBegin"); // evgen.rs:47
    void *ctx = nullptr;
    rc = tp_create(device, &ctx);
    if (rc != TP_OK) {
        std::fprintf(stderr, "Error: no usable HIP device %d (status %d)\n", device, rc);
        return 1;
    }
    // ### SIMULATION EXECUTION ### (main.rs:103-131, scheduling/mod.rs:31-59)
    union {
        tp_ranf_state ranf;
        tp_xoshiro_state xo;
        tp_xoshiro128_state xo128;
    } state;
    tp_rng_init(rng, &state);
    const uint64_t nb = (cfg.num_events + TP_BATCH_SIZE - 1) / TP_BATCH_SIZE;
    std::vector<tp_acc> batches(nb);
    tp_acc merged;
    rc = tp_run(ctx, &params, rng, seeding, &state, cfg.num_events, batches.data(), &merged, nullptr);
    if (rc != TP_OK) {
        std::fprintf(stderr, "Error: event loop failed: %s\n", tp_last_error(ctx));
        tp_destroy(ctx);
        return 1;
    }
    tp_final res;
    (precision == TP_PREC_F32 ? tp_finalize_f32 : tp_finalize)(&cfg, &merged, &res);
    // ### RESULTS DISPLAY AND STORAGE ### (main.rs:136-141)
    double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    rc = (precision == TP_PREC_F32 ? tp_dump_results_f32 : tp_dump_results)(&cfg, &res, elapsed, nullptr);
    tp_destroy(ctx);
    if (rc != TP_OK) {
        std::fprintf(stderr, "Error: failed to output the results\n");
        return 1;
    }
    return 0;
}
