// tp_host.cpp -- host side of the 3photons event loop: configuration,
// constants, finalisation and the reference's output formats.
//
// Restates src/config.rs, src/coupling.rs, src/evgen.rs:40-77,
// src/resacc.rs:59-223, src/resfin.rs and src/output.rs of the reference
// (HadrienG2/3photons-rust).  Built with -ffp-contract=off: every expression
// keeps Rust's left-to-right evaluation and IEEE rounding; powi(k) with a
// run-time base is binary powering (x*x, (x*x)*(x*x), x*(x*x)) and powi on a
// compile-time constant base is the host pow(), which is what LLVM folds it to.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "tp_guard.h"
#include "tp_host.h"

namespace {

constexpr double PI = 3.14159265358979323846;
const double SQRT_2 = 1.41421356237309504880;

// ---- Rust-compatible number formatting -----------------------------------

// Display (`{}`): shortest digits that round-trip (to f64, or to f32 for
// f32 values), never an exponent.  The reference's shipped outputs print
// negative zero as "0".
template <class T> bool round_trips(const char *buf, T x);
template <> bool round_trips<double>(const char *buf, double x) { return std::strtod(buf, nullptr) == x; }
template <> bool round_trips<float>(const char *buf, float x) { return std::strtof(buf, nullptr) == x; }

template <class T> std::string rust_display_t(T x)
{
    if (std::isnan(x))
        return "NaN";
    if (std::isinf(x))
        return x > 0 ? "inf" : "-inf";
    if (x == 0)
        return "0";
    char buf[64];
    int p;
    const int maxp = sizeof(T) == 4 ? 9 : 17;
    for (p = 1; p <= maxp; ++p) {
        std::snprintf(buf, sizeof buf, "%.*e", p - 1, (double)x);
        if (round_trips<T>(buf, x))
            break;
    }
    // buf = [-]d[.ddd]e[+-]XX
    std::string s(buf);
    bool neg = s[0] == '-';
    if (neg)
        s.erase(0, 1);
    size_t epos = s.find('e');
    int exp10 = std::atoi(s.c_str() + epos + 1);
    std::string digits;
    for (size_t i = 0; i < epos; ++i)
        if (std::isdigit(static_cast<unsigned char>(s[i])))
            digits += s[i];
    while (digits.size() > 1 && digits.back() == '0')
        digits.pop_back();
    std::string out;
    if (exp10 >= 0) {
        if (static_cast<int>(digits.size()) <= exp10 + 1) {
            out = digits + std::string(static_cast<size_t>(exp10 + 1) - digits.size(), '0');
        } else {
            out = digits.substr(0, static_cast<size_t>(exp10 + 1)) + "." +
                  digits.substr(static_cast<size_t>(exp10 + 1));
        }
    } else {
        out = "0." + std::string(static_cast<size_t>(-exp10 - 1), '0') + digits;
    }
    return neg ? "-" + out : out;
}

// `{:.Ne}` for f64: exact decimal rounding like printf, Rust's exponent
// spelling (no '+', no padding), zero printed unsigned as in the shipped
// outputs.
std::string rust_exp(double x, int prec)
{
    if (std::isnan(x))
        return "NaN";
    if (std::isinf(x))
        return x > 0 ? "inf" : "-inf";
    if (x == 0)
        x = 0.0;
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.*e", prec, x);
    std::string s(buf);
    size_t epos = s.find('e');
    int e = std::atoi(s.c_str() + epos + 1);
    return s.substr(0, epos) + "e" + std::to_string(e);
}

// `{:.N}` for f64.
std::string rust_fixed(double x, int prec)
{
    if (std::isnan(x))
        return "NaN";
    if (std::isinf(x))
        return x > 0 ? "inf" : "-inf";
    char buf[512];
    std::snprintf(buf, sizeof buf, "%.*f", prec, x);
    return buf;
}

std::string pad_left(const std::string &s, size_t w)
{
    return s.size() >= w ? s : std::string(w - s.size(), ' ') + s;
}
std::string pad_right(const std::string &s, size_t w)
{
    return s.size() >= w ? s : s + std::string(w - s.size(), ' ');
}

// write_engineering (output.rs:230-269), %g-like with SIG_DIGITS significant
// digits, the logarithm taken in the value's own type.
template <class T> std::string write_engineering_t(T x, int sig_digits)
{
    int precision = sig_digits - 1;
    if (x == 0)
        return "0";
    T log_x = std::log10(std::fabs(x));
    if (log_x >= (T)-3. && log_x <= static_cast<T>(sig_digits)) {
        long p = static_cast<long>(precision) - static_cast<long>(std::trunc(log_x));
        if (log_x < (T)0.)
            p += 1;
        if (p < 0)
            p = 0;
        std::string s = rust_fixed((double)x, static_cast<int>(p));
        if (s.find('.') != std::string::npos) {
            while (!s.empty() && s.back() == '0')
                s.pop_back();
            if (!s.empty() && s.back() == '.')
                s.pop_back();
        }
        return s;
    }
    return rust_exp((double)x, precision);
}

// f64::DIGITS - 1 or f32::DIGITS - 1 (output.rs:26)
template <class T> constexpr int sig_digits() { return sizeof(T) == 4 ? 6 - 1 : 15 - 1; }

struct Out {
    std::string s;
    void line(const std::string &l) { s += l + "\n"; }
    // writeln_3p with a (key, value) pair (output.rs:180-227)
    void kv(const char *key, const std::string &v) { s += " " + pad_right(key, 31) + ": " + v + "\n"; }
};

int copy_out(const std::string &s, char *buf, int len)
{
    if (buf && len > 0) {
        size_t n = s.size() < static_cast<size_t>(len - 1) ? s.size() : static_cast<size_t>(len - 1);
        std::memcpy(buf, s.data(), n);
        buf[n] = 0;
    }
    return static_cast<int>(s.size()) + 1;
}

// Rust's f64::from_str accepts decimal floats and inf/nan; strtod also takes
// hex floats, which Rust rejects.
bool parse_f64(const std::string &t, double *out)
{
    if (t.empty())
        return false;
    if (t.find('x') != std::string::npos || t.find('X') != std::string::npos)
        return false;
    errno = 0;
    char *end = nullptr;
    double v = std::strtod(t.c_str(), &end);
    if (*end != 0)
        return false;
    *out = v;
    return true;
}

bool parse_u64(const std::string &t, uint64_t *out)
{
    size_t i = (t.size() > 1 && t[0] == '+') ? 1 : 0;
    if (i >= t.size())
        return false;
    uint64_t v = 0;
    for (; i < t.size(); ++i) {
        if (!std::isdigit(static_cast<unsigned char>(t[i])))
            return false;
        uint64_t d = static_cast<uint64_t>(t[i] - '0');
        if (v > (UINT64_MAX - d) / 10)
            return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

bool parse_i32(const std::string &t, int32_t *out)
{
    size_t i = 0;
    bool neg = false;
    if (!t.empty() && (t[0] == '+' || t[0] == '-')) {
        neg = t[0] == '-';
        i = 1;
    }
    if (i >= t.size())
        return false;
    long long v = 0;
    for (; i < t.size(); ++i) {
        if (!std::isdigit(static_cast<unsigned char>(t[i])))
            return false;
        v = v * 10 + (t[i] - '0');
        if (v > 2147483648LL)
            return false;
    }
    if (neg)
        v = -v;
    if (v > 2147483647LL || v < -2147483648LL)
        return false;
    *out = static_cast<int32_t>(v);
    return true;
}

// ConfigItem::parse_bool (config.rs:187-195): Fortran booleans, else Rust's.
bool parse_bool(const std::string &t, int32_t *out)
{
    std::string low;
    for (char c : t)
        low += static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    if (low == ".true." || t == "true") {
        *out = 1;
        return true;
    }
    if (low == ".false." || t == "false") {
        *out = 0;
        return true;
    }
    return false;
}

int cfg_error(char *err, int len, const std::string &msg)
{
    copy_out(msg, err, len);
    return TP_ERR_CONFIG;
}

} // namespace

extern "C" int tp_config_load(const char *path, tp_config *cfg, char *err, int err_len)
{
    return tp_guard(nullptr, [&]() -> int {
        static const char *names[18] = {"num_events", "e_total", "beam_photons_cut", "photon_photon_cut",
                                        "e_min", "beam_photon_plane_cut", "alpha", "alpha_z",
                                        "gev2_to_picobarn", "m_z0", "g_z0", "sin2_weinberg",
                                        "branching_ep_em", "beta_plus", "beta_moins", "num_bins",
                                        "impr", "plot"};
        if (!path || !cfg)
            return cfg_error(err, err_len, "null argument");
        std::memset(cfg, 0, sizeof *cfg);
        FILE *f = std::fopen(path, "rb");
        if (!f)
            return cfg_error(err, err_len, std::string("failed to load the configuration: cannot open ") + path);
        std::string text;
        char chunk[4096];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof chunk, f)) > 0)
            text.append(chunk, got);
        std::fclose(f);
        // str::lines() then split_whitespace().next() per line (config.rs:69-71).
        std::string toks[18];
        int n = 0;
        size_t pos = 0;
        while (n < 18 && pos <= text.size()) {
            size_t eol = text.find('\n', pos);
            std::string line = text.substr(pos, eol == std::string::npos ? std::string::npos : eol - pos);
            pos = (eol == std::string::npos) ? text.size() + 1 : eol + 1;
            size_t a = 0;
            while (a < line.size() && std::isspace(static_cast<unsigned char>(line[a])))
                ++a;
            if (a == line.size())
                continue;
            size_t b = a;
            while (b < line.size() && !std::isspace(static_cast<unsigned char>(line[b])))
                ++b;
            toks[n++] = line.substr(a, b - a);
        }
        if (n < 18)
            return cfg_error(err, err_len, std::string("missing configuration of ") + names[n]);
        tp_config c;
        std::memset(&c, 0, sizeof c);
        double *dst[14] = {&c.e_total, &c.cut_beam, &c.cut_pp, &c.e_min, &c.cut_plane, &c.alpha,
                           &c.alpha_z, &c.gev2_to_picobarn, &c.m_z0, &c.g_z0, &c.sin2_weinberg,
                           &c.branching_ep_em, &c.beta_plus, &c.beta_minus};
        if (!parse_u64(toks[0], &c.num_events))
            return cfg_error(err, err_len, std::string("could not parse configuration of ") + names[0]);
        for (int i = 0; i < 14; ++i)
            if (!parse_f64(toks[1 + i], dst[i]))
                return cfg_error(err, err_len, std::string("could not parse configuration of ") + names[1 + i]);
        if (!parse_i32(toks[15], &c.num_bins))
            return cfg_error(err, err_len, std::string("could not parse configuration of ") + names[15]);
        if (!parse_bool(toks[16], &c.impr))
            return cfg_error(err, err_len, std::string("could not parse configuration of ") + names[16]);
        if (!parse_bool(toks[17], &c.plot))
            return cfg_error(err, err_len, std::string("could not parse configuration of ") + names[17]);
        c.parsed = 1;
        *cfg = c;
        // The ensure! checks of config.rs:110-124 come after the echo.
        if (c.num_events == 0)
            return cfg_error(err, err_len, "Please simulate at least one event");
        if (c.plot)
            return cfg_error(err, err_len, "Plotting is not supported by this version");
        if (c.impr)
            return cfg_error(err, err_len,
                             "Individual result printing is not supported. This debugging feature has a "
                             "run-time performance cost even when unused. It should be implemented at "
                             "compile-time instead.");
        if (err && err_len > 0)
            err[0] = 0;
        return TP_OK;
    });
}

template <class T> int config_format(const tp_config *c, char *buf, int len)
{
    Out o;
    o.line("ITOT           : " + std::to_string(c->num_events));
    o.line("ETOT           : " + rust_display_t<T>((T)c->e_total));
    o.line("oCutpar.ACUT   : " + rust_display_t<T>((T)c->cut_beam));
    o.line("oCutpar.BCUT   : " + rust_display_t<T>((T)c->cut_pp));
    o.line("oCutpar.EMIN   : " + rust_display_t<T>((T)c->e_min));
    o.line("oCutpar.SINCUT : " + rust_display_t<T>((T)c->cut_plane));
    o.line("ALPHA          : " + rust_display_t<T>((T)c->alpha));
    o.line("ALPHAZ         : " + rust_display_t<T>((T)c->alpha_z));
    o.line("CONVERS        : " + rust_display_t<T>((T)c->gev2_to_picobarn));
    o.line("oParam.MZ0     : " + rust_display_t<T>((T)c->m_z0));
    o.line("oParam.GZ0     : " + rust_display_t<T>((T)c->g_z0));
    o.line("SIN2W          : " + rust_display_t<T>((T)c->sin2_weinberg));
    o.line("BREPEM         : " + rust_display_t<T>((T)c->branching_ep_em));
    o.line("BETAPLUS       : " + rust_display_t<T>((T)c->beta_plus));
    o.line("BETAMOINS      : " + rust_display_t<T>((T)c->beta_minus));
    o.line("NBIN           : " + std::to_string(c->num_bins));
    o.line(std::string("oParam.IMPR    : ") + (c->impr ? "true" : "false"));
    o.line(std::string("PLOT           : ") + (c->plot ? "true" : "false"));
    return copy_out(o.s, buf, len);
}

extern "C" int tp_config_format(const tp_config *c, char *buf, int len) { return tp_guard(nullptr, [&]() -> int { return config_format<double>(c, buf, len); }); }

extern "C" int tp_config_format_f32(const tp_config *c, char *buf, int len) { return tp_guard(nullptr, [&]() -> int { return config_format<float>(c, buf, len); }); }

namespace {

// Scalar type of a run: double, or float for the all-f32 mode, where the
// reference parses, computes and prints in f32 (numeric.rs:6-14).
template <class T> struct Real;
template <> struct Real<double> {
    static constexpr double PI_ = PI;
    static double cv(double x) { return x; }
    // (2 pi)^-5: the reference folds the powi at compile time with the host pow()
    static double pow_2pi_m5() { return std::pow(2. * PI, -5.); }
};
template <> struct Real<float> {
    static constexpr float PI_ = 3.14159265358979323846f;
    static float cv(double x) { return (float)x; }
    static float pow_2pi_m5() { return (float)std::pow((double)(2.f * PI_), -5.); }
};

// Constants of ResultsAccumulator::new (resacc.rs:59-100) and
// EventGenerator::new (evgen.rs:40-77).
template <class T> struct AccConsts {
    T ev_weight, fact_com, norm_weight, propagator, delta, relat_width;
    T sc[5];
};

template <class T> AccConsts<T> acc_consts(const tp_config *c)
{
    using R = Real<T>;
    const T e_total = R::cv(c->e_total), conv = R::cv(c->gev2_to_picobarn), g_z0 = R::cv(c->g_z0);
    const T m_z0 = R::cv(c->m_z0), sin2w = R::cv(c->sin2_weinberg);
    const T beta_plus = R::cv(c->beta_plus), beta_minus = R::cv(c->beta_minus);
    AccConsts<T> k;
    // evgen.rs:49-62, NUM_OUTGOING = 3.
    T z_n = (T)2. * std::log(R::PI_ / (T)2.);
    z_n -= (T)2. * std::log((T)1.);
    z_n = z_n - std::log((T)2.);
    T ln_weight = ((T)2. * (T)3. - (T)4.) * std::log(e_total) + z_n;
    k.ev_weight = std::exp(ln_weight);
    // resacc.rs:64-100.
    T fact_com = (T)1. / (T)6. * conv;
    T relat_width = g_z0 / m_z0;
    T p_aa = (T)2.;
    T p_ab = (T)1. - (T)4. * sin2w;
    T p_bb = p_ab + (T)8. * (sin2w * sin2w);
    T m2 = m_z0 * m_z0;
    T m4 = m2 * m2;
    T c_aa = fact_com * p_aa;
    T c_ab = fact_com * p_ab / m2;
    T c_bb = fact_com * p_bb / m4;
    T r = e_total / m_z0;
    T dzeta = r * r;
    T delta = (dzeta - (T)1.) / relat_width;
    T propagator = (T)1. / ((T)1. + delta * delta);
    T n_ev = static_cast<T>(c->num_events);
    T norm = R::pow_2pi_m5() / n_ev;
    T norm_weight = k.ev_weight * norm;
    T com = norm_weight / (T)4.;
    T aa = com * c_aa;
    T bb = com * c_bb * propagator / (relat_width * relat_width);
    T ab = com * c_ab * (T)2. * beta_plus * propagator / relat_width;
    k.fact_com = fact_com;
    k.norm_weight = norm_weight;
    k.propagator = propagator;
    k.delta = delta;
    k.relat_width = relat_width;
    k.sc[0] = aa;
    k.sc[1] = bb * (beta_plus * beta_plus);
    k.sc[2] = bb * (beta_minus * beta_minus);
    k.sc[3] = ab * delta;
    k.sc[4] = -ab;
    return k;
}

template <class T> void params_from_config(const tp_config *c, tp_params *p)
{
    using R = Real<T>;
    std::memset(p, 0, sizeof *p);
    AccConsts<T> k = acc_consts<T>(c);
    const T alpha = R::cv(c->alpha), alpha_z = R::cv(c->alpha_z), sin2w = R::cv(c->sin2_weinberg);
    const T m_z0 = R::cv(c->m_z0);
    // Couplings::new (coupling.rs:23-33).
    T e2 = (T)4. * R::PI_ * alpha;
    T e2_z = (T)4. * R::PI_ * alpha_z;
    T cos2w = (T)1. - sin2w;
    T m2 = m_z0 * m_z0;
    T m4 = m2 * m2;
    T g_beta = -std::sqrt(e2_z / ((T)4. * cos2w * sin2w)) / m4;
    T x = std::sqrt(e2);
    p->e_total = R::cv(c->e_total);
    p->cut_beam = R::cv(c->cut_beam);
    p->cut_pp = R::cv(c->cut_pp);
    p->e_min = R::cv(c->e_min);
    p->cut_plane = R::cv(c->cut_plane);
    p->g_a = -(x * (x * x));
    p->g_beta_p = g_beta;
    p->g_beta_m = g_beta;
    for (int i = 0; i < 5; ++i)
        p->sigma_contribs[i] = k.sc[i];
    p->ev_weight = k.ev_weight;
    p->sort_photons = 1;
    p->precision = sizeof(T) == 4 ? TP_PREC_F32 : TP_PREC_F64;
    p->faster_evgen = 0;
    p->reserved = 0;
}

template <class T> void acc_merge(tp_acc *into, const tp_acc *o)
{
    into->selected_events += o->selected_events;
    for (int i = 0; i < 5; ++i)
        into->spm2[i] = (T)into->spm2[i] + (T)o->spm2[i];
    for (int i = 0; i < 5; ++i)
        into->vars[i] = (T)into->vars[i] + (T)o->vars[i];
    into->sigma = (T)into->sigma + (T)o->sigma;
    into->variance = (T)into->variance + (T)o->variance;
}

// ResultsAccumulator::finalize (resacc.rs:142-223).
template <class T> void finalize(const tp_config *cfg, const tp_acc *acc, tp_final *out)
{
    using R = Real<T>;
    AccConsts<T> k = acc_consts<T>(cfg);
    const T e_total = R::cv(cfg->e_total), sin2w = R::cv(cfg->sin2_weinberg);
    const T g_z0 = R::cv(cfg->g_z0), m_z0 = R::cv(cfg->m_z0);
    T spm2v[5], varsv[5];
    T n_ev = static_cast<T>(cfg->num_events);
    for (int c = 0; c < 5; ++c) {
        const T s = (T)acc->spm2[c], q = (T)acc->vars[c];
        T v = (q - s * s / n_ev) / (n_ev - (T)1.);
        varsv[c] = std::sqrt(v / n_ev) / std::fabs(s / n_ev);
        spm2v[c] = s;
    }
    T spm2[2][5], vars[2][5];
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 5; ++c) {
            spm2[s][c] = spm2v[c];
            vars[s][c] = varsv[c];
        }
    T polar_p = (T)-2. * sin2w;
    T polar_m = (T)1. + polar_p;
    const T polars[2] = {polar_m, polar_p};
    for (int c = 1; c < 5; ++c)
        for (int s = 0; s < 2; ++s)
            spm2[s][c] *= polars[s];
    for (int c = 1; c < 3; ++c)
        for (int s = 0; s < 2; ++s)
            spm2[s][c] *= polars[s];
    T incident_flux = (T)1. / ((T)2. * (e_total * e_total));
    T scale = k.fact_com * incident_flux * k.norm_weight;
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 5; ++c)
            spm2[s][c] *= scale;
    T gm_z0 = g_z0 * m_z0;
    for (int c = 1; c < 5; ++c)
        for (int s = 0; s < 2; ++s)
            spm2[s][c] *= k.propagator / gm_z0;
    for (int c = 1; c < 3; ++c)
        for (int s = 0; s < 2; ++s)
            spm2[s][c] /= gm_z0;
    for (int s = 0; s < 2; ++s)
        spm2[s][3] *= k.delta;
    auto colsum = [&](int col) { return ((T)0. + spm2[0][col]) + spm2[1][col]; };
    auto inc_num = [&](int col) {
        T a = spm2[0][col] * vars[0][col], b = spm2[1][col] * vars[1][col];
        return std::sqrt((T)0. + (a * a + b * b));
    };
    T beta_min = std::sqrt(colsum(0) / colsum(1));
    T ss_denom = colsum(0);
    T ss_norm = (T)1. / ((T)2. * std::sqrt(ss_denom));
    T ss_p = colsum(1) * ss_norm;
    T ss_m = colsum(2) * ss_norm;
    T inc_common = inc_num(0) / ((T)2. * std::fabs(ss_denom));
    T inc_ss_p = inc_num(1) / std::fabs(colsum(1)) + inc_common;
    T inc_ss_m = inc_num(2) / std::fabs(colsum(2)) + inc_common;
    const T sig = (T)acc->sigma, var = (T)acc->variance;
    T variance = (var - sig * sig / n_ev) / (n_ev - (T)1.);
    T prec = std::sqrt(variance / n_ev) / std::fabs(sig / n_ev);
    T sigma = sig * incident_flux;
    out->selected_events = acc->selected_events;
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 5; ++c) {
            out->spm2[s][c] = spm2[s][c];
            out->vars[s][c] = vars[s][c];
        }
    out->sigma = sigma;
    out->prec = prec;
    out->variance = variance;
    out->beta_min = beta_min;
    out->ss_p = ss_p;
    out->inc_ss_p = inc_ss_p;
    out->ss_m = ss_m;
    out->inc_ss_m = inc_ss_m;
}

} // namespace

extern "C" int tp_params_from_config(const tp_config *c, tp_params *p)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!c || !p)
            return TP_ERR_ARG;
        params_from_config<double>(c, p);
        return TP_OK;
    });
}

extern "C" int tp_params_from_config_f32(const tp_config *c, tp_params *p)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!c || !p)
            return TP_ERR_ARG;
        params_from_config<float>(c, p);
        return TP_OK;
    });
}

extern "C" void tp_acc_merge(tp_acc *into, const tp_acc *o) { acc_merge<double>(into, o); }

extern "C" void tp_acc_merge_f32(tp_acc *into, const tp_acc *o) { acc_merge<float>(into, o); }

extern "C" int tp_finalize(const tp_config *cfg, const tp_acc *acc, tp_final *out)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!cfg || !acc || !out)
            return TP_ERR_ARG;
        finalize<double>(cfg, acc, out);
        return TP_OK;
    });
}

extern "C" int tp_finalize_f32(const tp_config *cfg, const tp_acc *acc, tp_final *out)
{
    return tp_guard(nullptr, [&]() -> int {
        if (!cfg || !acc || !out)
            return TP_ERR_ARG;
        finalize<float>(cfg, acc, out);
        return TP_OK;
    });
}

namespace {

// powi of a constant, which the compiler folds with the host pow() in f64
// and rounds to the value's type.
template <class T> T cpow(T x, int n) { return (T)std::pow((double)x, (double)n); }

// Text of res.data (output.rs:64-142) in the run's float type.
template <class T> std::string res_data_text(const tp_config *cfg, const tp_final *res)
{
    Out o;
    auto F = [](T x) { return write_engineering_t<T>(x, sig_digits<T>()); };
    auto C = [](double x) { return (T)x; };
    o.kv("Nombre d'evenements", std::to_string(cfg->num_events));
    o.kv("... apres coupure", std::to_string(res->selected_events));
    o.kv("energie dans le CdM      (GeV)", F(C(cfg->e_total)));
    o.kv("coupure / cos(photon,faisceau)", F(C(cfg->cut_beam)));
    o.kv("coupure / cos(photon,photon)", F(C(cfg->cut_pp)));
    o.kv("coupure / sin(normale,faisceau)", F(C(cfg->cut_plane)));
    o.kv("coupure sur l'energie    (GeV)", F(C(cfg->e_min)));
    o.kv("1/(constante de structure fine)", F((T)1. / C(cfg->alpha)));
    o.kv("1/(structure fine au pic)", F((T)1. / C(cfg->alpha_z)));
    o.kv("facteur de conversion GeV-2/pb", F(C(cfg->gev2_to_picobarn)));
    o.kv("Masse du Z0              (GeV)", F(C(cfg->m_z0)));
    o.kv("Largeur du Z0            (GeV)", F(C(cfg->g_z0)));
    o.kv("Sinus^2 Theta Weinberg", F(C(cfg->sin2_weinberg)));
    o.kv("Taux de branchement Z--->e+e-", F(C(cfg->branching_ep_em)));
    o.kv("Beta plus", F(C(cfg->beta_plus)));
    o.kv("Beta moins", F(C(cfg->beta_minus)));
    o.line(" ---------------------------------------------");
    o.kv("Section Efficace          (pb)", F(C(res->sigma)));
    o.kv("Ecart-Type                (pb)", F(C(res->sigma) * C(res->prec)));
    o.kv("Precision Relative", F(C(res->prec)));
    o.line(" ---------------------------------------------");
    o.kv("Beta minimum", F(C(res->beta_min)));
    o.kv("Stat. Significance  B+(pb-1/2)", F(C(res->ss_p)));
    o.kv("Incert. Stat. Sign. B+(pb-1/2)", F(C(res->ss_p) * C(res->inc_ss_p)));
    o.kv("Stat. Significance  B-(pb-1/2)", F(C(res->ss_m)));
    o.kv("Incert. Stat. Sign. B-(pb-1/2)", F(C(res->ss_m) * C(res->inc_ss_m)));
    o.s += "\n";
    const int decs = std::min(sig_digits<T>() - 1, 7); // clamp_max(SIG_DIGITS - 1, 7)
    const size_t width = static_cast<size_t>(decs) + 8;
    for (int sp = 0; sp < 2; ++sp) {
        for (int k = 0; k < 5; ++k) {
            const T v = C(res->spm2[sp][k]), r = C(res->vars[sp][k]);
            o.s += pad_left(std::to_string(sp + 1), 3) + pad_left(std::to_string(k + 1), 3) +
                   pad_left(rust_exp((double)v, decs), width) + pad_left(rust_exp((double)(std::fabs(v) * r), decs), width) +
                   pad_left(rust_exp((double)r, decs), width) + "\n";
        }
        o.s += "\n";
    }
    for (int k = 0; k < 5; ++k) {
        const T s0 = C(res->spm2[0][k]), s1 = C(res->spm2[1][k]);
        T tmp1 = ((T)0. + s0) + s1;
        T a = s0 * C(res->vars[0][k]), b = s1 * C(res->vars[1][k]);
        T tmp2 = std::sqrt((T)0. + (a * a + b * b));
        o.s += "   " + pad_left(std::to_string(k + 1), 3) + pad_left(rust_exp((double)(tmp1 / (T)4.), decs), width) +
               pad_left(rust_exp((double)(tmp2 / (T)4.), decs), width) +
               pad_left(rust_exp((double)(tmp2 / std::fabs(tmp1)), decs), width) + "\n";
    }
    return o.s;
}

// FinalResults::eric + fawzi (resfin.rs:66-194) in the run's float type.
template <class T> std::string report_text(const tp_config *cfg, const tp_final *res)
{
    Out o;
    const T PI_T = (T)PI;
    auto C = [](double x) { return (T)x; };
    T spm2[2][5], vars[2][5];
    for (int s = 0; s < 2; ++s)
        for (int c = 0; c < 5; ++c) {
            spm2[s][c] = C(res->spm2[s][c]);
            vars[s][c] = C(res->vars[s][c]);
        }
    const T m_z0 = C(cfg->m_z0), g_z0 = C(cfg->g_z0), conv = C(cfg->gev2_to_picobarn);
    const T br = C(cfg->branching_ep_em), e = C(cfg->e_total);
    auto D = [](T x) { return (double)x; };
    // FinalResults::eric (resfin.rs:66-97).
    T mu_th = br * conv / ((T)8. * (T)9. * (T)5. * cpow<T>(PI_T, 2) * m_z0 * g_z0);
    T sigma0[2], alpha0[2], beta0[2], lambda0[2], mu0[2];
    for (int s = 0; s < 2; ++s) {
        sigma0[s] = spm2[s][0] / (T)2.;
        alpha0[s] = spm2[s][4] / (T)2.;
        beta0[s] = -spm2[s][3] / (T)2.;
        lambda0[s] = (spm2[s][2] - spm2[s][1]) / (T)2.;
        mu0[s] = (spm2[s][2] + spm2[s][1]) / (T)2.;
    }
    T mu_num = (((((T)0. + spm2[0][1]) + spm2[1][1]) + spm2[0][2]) + spm2[1][2]) / (T)4.;
    o.s += "\n";
    o.line("       :        -          +");
    o.line("sigma0  : " + rust_fixed(D(sigma0[0]), 6) + " | " + rust_fixed(D(sigma0[1]), 6));
    o.line("alpha0  : " + rust_exp(D(alpha0[0]), 5) + " | " + rust_exp(D(alpha0[1]), 4));
    o.line("beta0   : " + rust_display_t<T>(beta0[0]) + " | " + rust_display_t<T>(beta0[1]));
    o.line("lambda0 : " + rust_fixed(D(lambda0[0]), 4) + " | " + rust_fixed(D(lambda0[1]), 4));
    o.line("mu0     : " + rust_fixed(D(mu0[0]), 4) + " | " + rust_fixed(D(mu0[1]), 5));
    o.line("mu/lamb : " + rust_fixed(D(mu0[0] / lambda0[0]), 5) + " | " + rust_fixed(D(mu0[1] / lambda0[1]), 5));
    o.line("mu (num): " + rust_fixed(D(mu_num), 4));
    o.line("rapport : " + rust_fixed(D(mu_num / mu_th), 6));
    o.line("mu (th) : " + rust_fixed(D(mu_th), 4));
    // FinalResults::fawzi (resfin.rs:101-194).
    T mre = m_z0 / e;
    T gre = g_z0 * m_z0 / (e * e);
    T x = (T)1. - mre * mre;
    T den = x * x + gre * gre;
    T sdz_re = x / den, sdz_im = -gre / den;
    T del = ((T)1. - C(cfg->cut_pp)) / (T)2.;
    T eps = (T)2. * C(cfg->e_min) / e;
    T bra = m_z0 / ((T)3. * (T)6. * cpow<T>(PI_T, 3) * (T)16. * (T)120.);
    T r = e / m_z0;
    T r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    T sig = (T)12. * PI_T / (m_z0 * m_z0) * br * g_z0 * bra / (e * e) * r8 * (sdz_re * sdz_re + sdz_im * sdz_im) * conv;
    T eps_2 = eps * eps;
    T eps_4 = eps_2 * eps_2;
    T del_2 = del * del;
    T del_3 = del * del_2;
    T f1 = (T)1. - (T)15. * eps_4 - (T)9. / (T)7. * ((T)1. - (T)70. * eps_4) * del_2 +
           (T)6. / (T)7. * ((T)1. + (T)70. * eps_4) * del_3;
    T g1 = (T)1. - (T)30. * eps_4 - (T)9. / (T)7. * ((T)1. - (T)70. * eps_4) * del - (T)90. * eps_4 * del_2 -
           (T)1. / (T)7. * ((T)1. - (T)420. * eps_4) * del_3;
    T g2 = (T)1. - (T)25. * eps_4 - (T)6. / (T)7. * ((T)1. - (T)70. * eps_4) * del -
           (T)3. / (T)7. * ((T)1. + (T)210. * eps_4) * del_2 - (T)8. / (T)21. * ((T)1. - (T)52.5 * eps_4) * del_3;
    T g3 = (T)1. - (T)195. / (T)11. * eps_4 - (T)18. / (T)77. * ((T)1. - (T)7. * eps_4) * del -
           (T)9. / (T)11. * ((T)9. / (T)7. - (T)70. * eps_4) * del_2 -
           (T)8. / (T)11. * ((T)1. - (T)105. / (T)11. * eps_4) * del_3;
    T c = C(cfg->cut_plane);
    T c3 = c * (c * c);
    T ff = f1 * ((T)1. - c3);
    T gg = g1 - (T)27. / (T)16. * g2 * c + (T)11. / (T)16. * g3 * c3;
    T sig_p = sig * (ff + (T)2. * gg);
    T sig_m = sig_p + (T)2. * sig * gg;
    T mc_p = (((T)0. + spm2[0][1]) + spm2[1][1]) / (T)4.;
    T mc_m = (((T)0. + spm2[0][2]) + spm2[1][2]) / (T)4.;
    auto incr = [&](int col) {
        T a = spm2[0][col] * vars[0][col], b = spm2[1][col] * vars[1][col];
        return std::sqrt((T)0. + (a * a + b * b)) / std::fabs(((T)0. + spm2[0][col]) + spm2[1][col]);
    };
    T incr_p = incr(1), incr_m = incr(2);
    o.s += "\n";
    o.line("s (pb) :   Sig_cut_Th    Sig_Th      Rapport");
    o.line("       :   Sig_Num");
    o.line("       :   Ecart_relatif  Incertitude");
    o.s += "\n";
    o.line("s+(pb) : " + rust_fixed(D(sig_p), 5) + " | " + rust_fixed(D(sig * (T)3.), 5) + " | " +
           rust_fixed(D(sig_p / ((T)3. * sig)), 6));
    o.line("       : " + rust_fixed(D(mc_p), 5));
    o.line("       : " + rust_fixed(D(mc_p / sig_p - (T)1.), 6) + " | " + rust_fixed(D(incr_p), 8) + " | " +
           rust_fixed(D((mc_p / sig_p - (T)1.) / incr_p), 2));
    o.s += "\n";
    o.line("s-(pb) : " + rust_fixed(D(sig_m), 5) + " | " + rust_fixed(D(sig * (T)5.), 4) + " | " +
           rust_fixed(D(sig_m / ((T)5. * sig)), 6));
    o.line("       : " + rust_fixed(D(mc_m), 5));
    o.line("       : " + rust_fixed(D(mc_m / sig_m - (T)1.), 6) + " | " + rust_fixed(D(incr_m), 9) + " | " +
           rust_fixed(D((mc_m / sig_m - (T)1.) / incr_m), 2));
    o.s += "\n";
    return o.s;
}

} // namespace

extern "C" int tp_format_res_data(const tp_config *cfg, const tp_final *res, char *buf, int len)
{
    return tp_guard(nullptr, [&]() -> int {
        return copy_out(res_data_text<double>(cfg, res), buf, len);
    });
}

extern "C" int tp_format_res_data_f32(const tp_config *cfg, const tp_final *res, char *buf, int len)
{
    return tp_guard(nullptr, [&]() -> int {
        return copy_out(res_data_text<float>(cfg, res), buf, len);
    });
}

extern "C" int tp_format_report(const tp_config *cfg, const tp_final *res, char *buf, int len)
{
    return tp_guard(nullptr, [&]() -> int {
        return copy_out(report_text<double>(cfg, res), buf, len);
    });
}

extern "C" int tp_format_report_f32(const tp_config *cfg, const tp_final *res, char *buf, int len)
{
    return tp_guard(nullptr, [&]() -> int {
        return copy_out(report_text<float>(cfg, res), buf, len);
    });
}

template <class T> int dump_results(const tp_config *cfg, const tp_final *res, double elapsed_s, const char *dir)
{
    if (!cfg || !res)
        return TP_ERR_ARG;
    const std::string report = report_text<T>(cfg, res);
    std::fputs(report.c_str(), stdout);
    std::fflush(stdout);
    // "[day]-[month repr:short]-[year repr:last_two] [hour]:[minute]:[second]" (output.rs:37-39), UTC.
    std::time_t now = std::time(nullptr);
    std::tm tmv;
    gmtime_r(&now, &tmv);
    char stamp[64];
    std::strftime(stamp, sizeof stamp, "%d-%b-%y %H:%M:%S", &tmv);
    std::string base = dir ? std::string(dir) + "/" : std::string();
    {
        FILE *f = std::fopen((base + "res.times").c_str(), "w");
        if (!f)
            return TP_ERR_IO;
        Out o;
        o.s += std::string(" ") + stamp + "\n";
        o.line(" ---------------------------------------------");
        o.kv("Temps ecoule", "???");
        o.kv("Temps ecoule utilisateur", write_engineering_t<T>((T)elapsed_s, sig_digits<T>()));
        o.kv("Temps ecoule systeme", "???");
        o.kv("Temps ecoule par evenement",
             write_engineering_t<T>((T)elapsed_s / static_cast<T>(cfg->num_events), sig_digits<T>()));
        std::fputs(o.s.c_str(), f);
        std::fclose(f);
    }
    {
        const std::string text = res_data_text<T>(cfg, res);
        FILE *f = std::fopen((base + "res.data").c_str(), "w");
        if (!f)
            return TP_ERR_IO;
        std::fputs(text.c_str(), f);
        std::fclose(f);
    }
    {
        FILE *f = std::fopen((base + "pil.mc").c_str(), "a");
        if (!f)
            return TP_ERR_IO;
        auto colsum = [&](int col) { return ((T)0. + (T)res->spm2[0][col]) + (T)res->spm2[1][col]; };
        const T bp = (T)cfg->beta_plus, bm = (T)cfg->beta_minus;
        T res1 = colsum(0);
        T res2 = colsum(1) * (bp * bp);
        T res3 = colsum(2) * (bm * bm);
        T res4 = colsum(3) * bp;
        auto Dp = [](T x) { return rust_display_t<T>(x); };
        std::string l = Dp((T)cfg->e_total) + " " + Dp(res1 / (T)4.) + " " + Dp(res2 / (T)4.) + " " +
                        Dp(res3 / (T)4.) + " " + Dp(res4 / (T)4.) + " " + Dp((res1 + res2 + res3 + res4) / (T)4.) +
                        " " + Dp((T)res->sigma);
        std::fprintf(f, "%s\n%s\n", stamp, l.c_str());
        std::fclose(f);
    }
    return TP_OK;
}

extern "C" int tp_dump_results(const tp_config *cfg, const tp_final *res, double elapsed_s, const char *dir)
{
    return tp_guard(nullptr, [&]() -> int {
        return dump_results<double>(cfg, res, elapsed_s, dir);
    });
}

extern "C" int tp_dump_results_f32(const tp_config *cfg, const tp_final *res, double elapsed_s, const char *dir)
{
    return tp_guard(nullptr, [&]() -> int {
        return dump_results<float>(cfg, res, elapsed_s, dir);
    });
}
