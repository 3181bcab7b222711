// deck.cpp -- the text input deck loader (host only, no GPU code).
//
// Reference: Deck / load_deck (include/pic/deck.hpp:31-274).  INI sections,
// `key = value` lines, `#` / `!` comments, repeated `[species <name>]` blocks;
// unknown sections and keys are ignored.  Numbers follow the reference's
// std::stod / std::stol (leading number, trailing text ignored); the
// finalize step turns physical inputs into code-unit RunParams.
#include "aw.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <numbers>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace aw {
void set_last_error(const std::string& msg);
}

namespace {

struct DeckError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

std::string trim(const std::string& s) {
    const char* ws = " \t\r\n";
    const auto a = s.find_first_not_of(ws);
    if (a == std::string::npos) return "";
    return s.substr(a, s.find_last_not_of(ws) - a + 1);
}

bool as_bool(const std::string& v) {
    return v == "1" || v == "true" || v == "yes" || v == "on" || v == ".true.";
}

// up to three whitespace-separated numbers; missing entries stay 0
void as_vec3(const std::string& v, double out[3]) {
    out[0] = out[1] = out[2] = 0.0;
    std::istringstream is(v);
    for (int i = 0; i < 3; ++i)
        if (!(is >> out[i])) {
            out[i] = 0.0;
            break;
        }
}

void copy_str(char* dst, size_t cap, const std::string& s) {
    std::snprintf(dst, cap, "%s", s.c_str());
}

void append_word(char* dst, size_t cap, const std::string& w) {
    std::string cur(dst);
    if (!cur.empty()) cur += ' ';
    cur += w;
    copy_str(dst, cap, cur);
}

class Parser {
public:
    explicit Parser(aw_deck& d) : d_(d) {}

    void line(std::string text, int no) {
        no_ = no;
        for (char c : {'#', '!'}) {
            const auto p = text.find(c);
            if (p != std::string::npos) text.resize(p);
        }
        text = trim(text);
        if (text.empty()) return;
        if (text.front() == '[') {
            const auto e = text.find(']');
            if (e == std::string::npos)
                throw DeckError("deck: bad section at line " + std::to_string(no));
            section_ = trim(text.substr(1, e - 1));
            if (section_.rfind("species", 0) == 0) add_species(trim(section_.substr(7)));
            return;
        }
        const auto eq = text.find('=');
        if (eq == std::string::npos)
            throw DeckError("deck: expected key=value at line " + std::to_string(no));
        key_ = trim(text.substr(0, eq));
        val_ = trim(text.substr(eq + 1));
        dispatch();
    }

private:
    aw_deck& d_;
    std::string section_, key_, val_;
    int no_ = 0;

    double num() const {
        try {
            return std::stod(val_);
        } catch (const std::exception&) {
            throw DeckError("deck: bad number '" + val_ + "' for " + key_ + " at line " +
                            std::to_string(no_));
        }
    }
    long inum() const {
        try {
            return std::stol(val_);
        } catch (const std::exception&) {
            throw DeckError("deck: bad integer '" + val_ + "' for " + key_ + " at line " +
                            std::to_string(no_));
        }
    }
    bool is(const char* k) const { return key_ == k; }

    void add_species(const std::string& name) {
        if (d_.nspecies >= AW_DECK_MAX_SPECIES)
            throw DeckError("deck: more than " + std::to_string(AW_DECK_MAX_SPECIES) +
                            " species blocks");
        aw_species& sp = d_.species[d_.nspecies++];
        aw_species_defaults(&sp);
        if (!name.empty()) copy_str(sp.name, sizeof(sp.name), name);
    }

    void dispatch() {
        aw_run_params& rp = d_.rp;
        const std::string& s = section_;
        if (s == "grid") {
            if (is("nx")) d_.nx = (int32_t)inum();
            else if (is("ny")) d_.ny = (int32_t)inum();
            else if (is("Lx")) d_.Lx = num();
            else if (is("Ly")) d_.Ly = num();
        } else if (s == "time") {
            if (is("dt")) rp.dt = num();
            else if (is("nsteps")) rp.nsteps = inum();
            else if (is("dump_every")) d_.dump_every = inum();
        } else if (s == "plasma") {
            if (is("qm")) rp.qm = num();
            else if (is("eps0")) rp.eps0 = num();
            else if (is("wpe")) rp.wpe = num();
            else if (is("noisy")) rp.noisy_load = as_bool(val_);
            else if (is("seed")) rp.rng_seed = (uint64_t)(unsigned long)inum();
            else if (is("cold_nc")) rp.cold_nc = num();
            else if (is("rel")) rp.rel = as_bool(val_) ? 1 : 0;
            else if (is("bnoise")) d_.bnoise = num();
            else if (is("outdir")) copy_str(d_.outdir, sizeof(d_.outdir), val_);
        } else if (s.rfind("species", 0) == 0) {
            if (d_.nspecies == 0) throw DeckError("deck: species key outside a species block");
            aw_species& sp = d_.species[d_.nspecies - 1];
            if (is("density")) sp.density = num();
            else if (is("ppc")) sp.ppc = (int32_t)inum();
            else if (is("uth")) as_vec3(val_, sp.uth);
            else if (is("ufl")) as_vec3(val_, sp.ufl);
            else if (is("rep")) sp.deltaf = val_ == "deltaf";
            else if (is("dist")) {
                if (val_ == "losscone") sp.dist = 1;
                else if (val_ != "bimax")
                    throw DeckError("deck: species dist must be bimax|losscone");
            } else if (is("kappa")) sp.lc_kappa = num();
            else if (is("rho")) sp.lc_rho = num();
            else if (is("taud")) sp.taud = num();
            else if (is("wdnoise")) sp.wdnoise = num();
        } else if (s == "field") {
            if (is("model")) {
                d_.darwin = val_ == "darwin";
                d_.yee = val_ == "yee";
            } else if (is("dx_wpe_c")) d_.dx_wpe_c = num();
            else if (is("c")) d_.c_direct = num();
            else if (is("ndc")) rp.ndc = (int32_t)inum();
            else if (is("jfilter")) rp.jfilter = (int32_t)inum();
            else if (is("tile_sort")) rp.tile_sort = (int32_t)inum();
        } else if (s == "antenna") {
            if (is("amp")) rp.ant_amp = num();
            else if (is("x0")) rp.ant_x0 = num();
            else if (is("sigma")) rp.ant_sigma = num();
            else if (is("w0")) rp.ant_w0 = num();
            else if (is("trmp")) rp.ant_trmp = num();
            else if (is("toff")) rp.ant_toff = num();
        } else if (s == "boundary") {
            if (is("x")) rp.bnd_x = val_ == "hybrid" ? 2 : val_ == "damping" ? 1 : 0;
            else if (is("nd")) rp.bnd_nd = (int32_t)inum();
            else if (is("numax")) rp.bnd_numax = num();
        } else if (s == "background") {
            if (is("wce")) {
                d_.wce = num();
                rp.wce = d_.wce;
            } else if (is("theta_deg")) d_.theta_deg = num();
            else if (is("B0")) {
                double v[3];
                as_vec3(val_, v);
                for (int i = 0; i < 3; ++i) rp.B0[i] = (float)v[i];
                d_.b0_direct = 1;
            } else if (is("profile")) {
                if (val_ == "parabolic") rp.b0_prof = 1;
                else if (val_ == "dipole") rp.b0_prof = 2;
                else if (val_ != "uniform")
                    throw DeckError(
                        "deck: [background] profile must be uniform|parabolic|dipole");
            } else if (is("a")) rp.b0_a = num();
            else if (is("lre")) rp.b0_lre = num();
            else if (is("xc")) {
                rp.b0_xc = num();
                d_.b0_xc_set = 1;
            }
        } else if (s == "pump") {
            if (is("enable")) d_.pump_enable = as_bool(val_);
            else if (is("mode")) d_.pump_M = (int32_t)inum();
            else if (is("w0")) d_.pump_w0 = num();
            else if (is("ex0")) d_.pump_ex0 = num();
            else if (is("ey0")) d_.pump_ey0 = num();
            else if (is("ez0")) d_.pump_ez0 = num();
            else if (is("amp")) d_.pump_amp = num();
            else if (is("trmp")) d_.pump_trmp = num();
            else if (is("toff")) d_.pump_toff = num();
        } else if (s == "diagnostics") {
            if (is("enable")) {
                std::istringstream in(val_);
                std::string m;
                while (in >> m) append_word(d_.diag_enable, sizeof(d_.diag_enable), m);
            } else if (is("tsnap")) d_.tsnap = num();
            else if (is("band_lo")) d_.band_lo = (int32_t)inum();
            else if (is("band_hi")) d_.band_hi = (int32_t)inum();
            else if (is("kt_stride")) d_.kt_stride = (int32_t)inum();
            else if (is("n_frames")) d_.n_frames = (int32_t)inum();
            else if (is("prefix")) copy_str(d_.prefix, sizeof(d_.prefix), val_);
            else if (is("probes")) {
                std::istringstream in(val_);
                double p;
                while (in >> p) {
                    if (d_.nprobes >= AW_DECK_MAX_PROBES)
                        throw DeckError("deck: more than " + std::to_string(AW_DECK_MAX_PROBES) +
                                        " probes");
                    d_.probes[d_.nprobes++] = p;
                }
            }
        }
    }
};

void deck_defaults(aw_deck& d) {
    std::memset(&d, 0, sizeof(d));
    aw_run_params_defaults(&d.rp);
    d.rp.eps0 = 1.0;
    d.rp.n0 = 1.0;
    d.rp.qm = -1.0;
    d.rp.wpe = 1.0;
    copy_str(d.outdir, sizeof(d.outdir), "deck_frames");
    d.pump_ex0 = 1.39;
    d.pump_ey0 = 1.81;
    d.pump_ez0 = -1.81;
    d.pump_amp = 1.0;
    d.kt_stride = 5;
    d.n_frames = 150;
}

// code-unit RunParams from the physical inputs (needs dx = Lx/nx)
// ---- the dipole background (include/pic/background_b0.hpp:57-146) --------------
// Along a dipole field line of equatorial distance L R_E, at magnetic
// latitude lam, B/B_eq = sqrt(1 + 3 sin^2 lam) / cos^6 lam and the arc length
// from the equator is s = L R_E (sin lam sqrt(1 + 3 sin^2 lam) +
// asinh(sqrt(3) sin lam)/sqrt(3)) / 2.  The Yee push evaluates B0(s) as an
// even polynomial in u = (s/s_ref)^2 of degree 5, least-squares fitted to 400
// midpoint samples of |s| <= s_ref and checked to 2e-4 relative.

double dipole_arc(double lam) {
    const double t = std::sin(lam);
    return 0.5 * (t * std::sqrt(1.0 + 3.0 * t * t) + std::asinh(std::sqrt(3.0) * t) / std::sqrt(3.0));
}

double dipole_b_over_beq(double s, double lre) {
    const double target = std::fabs(s) / lre;
    if (target <= 0.0) return 1.0;
    double lo = 0.0, hi = 0.5 * std::numbers::pi;
    if (target >= dipole_arc(hi)) throw DeckError("dipole: s beyond the field-line apex");
    for (int it = 0; it < 200; ++it) {  // s(lam) is increasing: bisection
        const double mid = 0.5 * (lo + hi);
        (dipole_arc(mid) < target ? lo : hi) = mid;
    }
    const double lam = 0.5 * (lo + hi), t = std::sin(lam), c2 = 1.0 - t * t;
    return std::sqrt(1.0 + 3.0 * t * t) / (c2 * c2 * c2);
}

void fit_dipole(aw_run_params& rp, double sref) {
    constexpr int K = 6, M = 400;
    rp.b0_sref = sref;
    double G[K][K] = {}, r[K] = {};  // normal equations of the sampled fit
    for (int m = 0; m < M; ++m) {
        const double sv = sref * (m + 0.5) / M, u = (sv / sref) * (sv / sref);
        const double f = dipole_b_over_beq(sv, rp.b0_lre);
        double pw[K];
        pw[0] = 1.0;
        for (int k = 1; k < K; ++k) pw[k] = pw[k - 1] * u;
        for (int a = 0; a < K; ++a) {
            r[a] += pw[a] * f;
            for (int b = 0; b < K; ++b) G[a][b] += pw[a] * pw[b];
        }
    }
    // Cholesky G = L L^T (G is symmetric positive definite), then two solves
    double Lc[K][K] = {};
    for (int a = 0; a < K; ++a) {
        for (int b = 0; b <= a; ++b) {
            double v = G[a][b];
            for (int k = 0; k < b; ++k) v -= Lc[a][k] * Lc[b][k];
            if (a == b) {
                if (!(v > 0.0)) throw DeckError("fit_dipole: normal matrix not positive definite");
                Lc[a][a] = std::sqrt(v);
            } else {
                Lc[a][b] = v / Lc[b][b];
            }
        }
    }
    double z[K];
    for (int a = 0; a < K; ++a) {
        double v = r[a];
        for (int k = 0; k < a; ++k) v -= Lc[a][k] * z[k];
        z[a] = v / Lc[a][a];
    }
    for (int a = K - 1; a >= 0; --a) {
        double v = z[a];
        for (int k = a + 1; k < K; ++k) v -= Lc[k][a] * rp.b0_c[k];
        rp.b0_c[a] = v / Lc[a][a];
    }
    double worst = 0.0;
    for (int m = 0; m <= 100; ++m) {
        const double sv = sref * m / 100.0, u = (sv / sref) * (sv / sref);
        double p = rp.b0_c[K - 1];
        for (int k = K - 2; k >= 0; --k) p = p * u + rp.b0_c[k];
        const double f = dipole_b_over_beq(sv, rp.b0_lre);
        worst = std::max(worst, std::fabs(p - f) / f);
    }
    if (worst > 2e-4)
        throw DeckError("fit_dipole: fit residual " + std::to_string(worst) + " exceeds 2e-4");
}

void finalize(aw_deck& d) {
    if (d.nx <= 0 || d.ny <= 0 || d.Lx <= 0.0 || d.Ly <= 0.0)
        throw DeckError("deck: [grid] nx/ny/Lx/Ly must all be set and positive");
    if (d.nspecies == 0) throw DeckError("deck: at least one [species ...] block is required");
    aw_run_params& rp = d.rp;
    const double dx = d.Lx / d.nx;
    constexpr double pi = std::numbers::pi;
    if (d.c_direct > 0.0) rp.c = d.c_direct;
    else if (d.dx_wpe_c > 0.0) rp.c = dx / d.dx_wpe_c;
    if (!d.b0_direct && d.wce != 0.0) {  // B0 at theta in the x-z plane
        const double th = d.theta_deg * pi / 180.0;
        rp.B0[0] = (float)(d.wce * std::cos(th));
        rp.B0[1] = 0.0f;
        rp.B0[2] = (float)(d.wce * std::sin(th));
    }
    if (rp.b0_prof) {
        if (!d.b0_xc_set) rp.b0_xc = 0.5 * d.Lx;
        if (rp.B0[1] != 0.f || rp.B0[2] != 0.f)
            throw DeckError(
                "deck: [background] profile=parabolic|dipole requires B0 along x (theta_deg = 0)");
        if (rp.b0_prof == 2) {
            if (rp.b0_lre <= 0.0)
                throw DeckError("deck: [background] profile=dipole requires lre = L*R_E > 0");
            // 1% past the farther wall, so the wall itself lies inside the fit
            fit_dipole(rp, 1.01 * std::max(rp.b0_xc, d.Lx - rp.b0_xc));
        }
    }
    if (d.pump_enable) {  // An et al. (2019) Table I amplitudes, 1e4 e E/(me wpe^2 dx) units
        const double s = d.pump_amp * dx / 1e4;
        rp.pump = 1;
        rp.pump_k0 = 2.0 * pi * d.pump_M / d.Lx;
        rp.pump_w0 = d.pump_w0;
        rp.pump_ex = d.pump_ex0 * s;
        rp.pump_ey = d.pump_ey0 * s;
        rp.pump_ez = d.pump_ez0 * s;
        rp.pump_trmp = d.pump_trmp;
        rp.pump_toff = d.pump_toff;
    }
}

}  // namespace

extern "C" aw_status aw_deck_load(const char* path, aw_deck* out) {
    try {
        if (!path || !out) throw DeckError("aw_deck_load: path/out is NULL");
        std::ifstream f(path);
        if (!f) throw DeckError(std::string("deck: cannot open ") + path);
        aw_deck d;
        deck_defaults(d);
        Parser p(d);
        std::string text;
        int no = 0;
        while (std::getline(f, text)) p.line(text, ++no);
        finalize(d);
        *out = d;
        return AW_OK;
    } catch (const std::exception& e) {
        aw::set_last_error(e.what());
        return AW_ERR_INVALID;
    }
}
