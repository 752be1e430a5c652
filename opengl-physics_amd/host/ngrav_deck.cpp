// ngrav_deck.cpp -- deck-driven batch runs of the gravity path on the GPU backend.
//
// The drop-in for `02_nbody_gravity --deck <file> --out <dir>` on the nbody_core path:
//   * the TOML deck keys of DeckSim::Configure (projects/nbody_core/src/DeckSim.cpp:26-74),
//     plus `solver.backend` ("gpu", the key the reference planned but never implemented,
//     N_Body_Gravity_Plan.md:113) and `solver.multipole` ("quadrupole" default -- System<3>
//     always adds quadrupoles, System.hpp:59 -- or "monopole");
//   * scenario defaults of BuildScenario (projects/nbody_core/src/Scenarios.cpp:22-330) for the
//     scenarios whose samplers exist here: plummer, hernquist, king, cluster, cold_collapse,
//     kepler, lagrange, ic_file;
//   * the frame loop of fw::RunHeadless (framework/src/HeadlessRunner.cpp:11-41): snapshot,
//     then `substeps_per_frame` leapfrog steps, `frames` times;
//   * the output layout of fw::OutputWriter (framework/src/OutputWriter.cpp:38-169): deck.toml
//     copy, diagnostics.csv ("t,step," + DeckSim::Info diagnostics, DeckSim.cpp:178-179),
//     frames/pos_NNNN.npy and vel_NNNN.npy (float32, N x 3), manifest.json.
// Energy, angular momentum, momentum and centre of mass come from the FP64 GPU diagnostics
// (System::TotalEnergy semantics, System.hpp:119-193) instead of an O(N^2) CPU loop.
// Softening kinds, the relative MAC, adaptive and block time steps run on the GPU backend;
// the FMM solvers and the scenarios without a sampler here raise std::runtime_error.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nbg.h"
#include "../../include/ngrav_host.h"

namespace fs = std::filesystem;

namespace {

// ---------------------------------------------------------------------------- TOML subset
// Tables ([a] / [a.b]), dotted keys, basic and literal strings, integers, floats, booleans
// and '#' comments -- the syntax the reference decks use (projects/02_nbody_gravity/decks).
class Deck {
public:
    static Deck FromFile(const std::string& path) {
        std::ifstream f(path, std::ios::binary);
        if (!f) throw std::runtime_error("deck: cannot open '" + path + "'");
        std::stringstream ss;
        ss << f.rdbuf();
        Deck d;
        d.source_ = ss.str();
        d.Parse();
        return d;
    }
    bool Has(const std::string& k) const { return kv_.count(k) != 0; }
    std::string GetString(const std::string& k, const std::string& def) const {
        auto it = kv_.find(k);
        return it == kv_.end() ? def : it->second.s;
    }
    double GetDouble(const std::string& k, double def) const {
        auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        if (it->second.type == 'n') return it->second.d;
        throw std::runtime_error("deck: '" + k + "' is not a number");
    }
    int GetInt(const std::string& k, int def) const { return static_cast<int>(std::llround(GetDouble(k, def))); }
    bool GetBool(const std::string& k, bool def) const {
        auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        if (it->second.type == 'b') return it->second.b;
        throw std::runtime_error("deck: '" + k + "' is not a boolean");
    }
    const std::string& Source() const { return source_; }

private:
    struct Value {
        char type = 's';
        std::string s;
        double d = 0.0;
        bool b = false;
    };
    static std::string Trim(const std::string& s) {
        size_t a = 0, b = s.size();
        while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
        while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
        return s.substr(a, b - a);
    }
    static std::string StripComment(const std::string& line) {
        bool inBasic = false, inLit = false;
        for (size_t i = 0; i < line.size(); ++i) {
            const char c = line[i];
            if (c == '"' && !inLit && (i == 0 || line[i - 1] != '\\')) inBasic = !inBasic;
            if (c == '\'' && !inBasic) inLit = !inLit;
            if (c == '#' && !inBasic && !inLit) return line.substr(0, i);
        }
        return line;
    }
    static std::string Key(std::string k) {
        k = Trim(k);
        std::string out;
        for (char c : k)
            if (!std::isspace(static_cast<unsigned char>(c))) out += c;
        return out;
    }
    void Parse() {
        std::istringstream in(source_);
        std::string line, table;
        int lineNo = 0;
        while (std::getline(in, line)) {
            ++lineNo;
            const std::string t = Trim(StripComment(line));
            if (t.empty()) continue;
            if (t.front() == '[') {
                if (t.back() != ']' || t.size() < 3) throw std::runtime_error("deck: bad table header at line " +
                                                                               std::to_string(lineNo));
                table = Key(t.substr(1, t.size() - 2));
                continue;
            }
            const size_t eq = t.find('=');
            if (eq == std::string::npos) throw std::runtime_error("deck: expected key = value at line " +
                                                                  std::to_string(lineNo));
            const std::string key = (table.empty() ? "" : table + ".") + Key(t.substr(0, eq));
            const std::string v = Trim(t.substr(eq + 1));
            Value val;
            if (v.size() >= 2 && v.front() == '"' && v.back() == '"') {
                for (size_t i = 1; i + 1 < v.size(); ++i) {
                    if (v[i] == '\\' && i + 2 < v.size()) {
                        const char e = v[++i];
                        val.s += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                    } else {
                        val.s += v[i];
                    }
                }
            } else if (v.size() >= 2 && v.front() == '\'' && v.back() == '\'') {
                val.s = v.substr(1, v.size() - 2);
            } else if (v == "true" || v == "false") {
                val.type = 'b';
                val.b = (v == "true");
            } else {
                std::string num;
                for (char c : v)
                    if (c != '_') num += c;
                char* end = nullptr;
                val.d = std::strtod(num.c_str(), &end);
                if (end == num.c_str() || *end != '\0')
                    throw std::runtime_error("deck: unsupported value '" + v + "' at line " + std::to_string(lineNo));
                val.type = 'n';
            }
            kv_[key] = val;
        }
    }
    std::string source_;
    std::map<std::string, Value> kv_;
};

// ---------------------------------------------------------------------------- output
void WriteNpyF32(const fs::path& path, const float* data, int rows, int cols) {
    std::ostringstream dict;
    dict << "{'descr': '<f4', 'fortran_order': False, 'shape': (" << rows << ", " << cols << "), }";
    std::string header = dict.str();
    const size_t pre = 10 + header.size() + 1;
    const size_t padded = (pre + 63) & ~size_t(63);
    header.append(padded - pre, ' ');
    header.push_back('\n');
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("output: cannot open " + path.string());
    const char magic[8] = {'\x93', 'N', 'U', 'M', 'P', 'Y', '\x01', '\x00'};
    f.write(magic, 8);
    const uint16_t hl = static_cast<uint16_t>(header.size());
    f.put(static_cast<char>(hl & 0xff)).put(static_cast<char>(hl >> 8));
    f.write(header.data(), static_cast<std::streamsize>(header.size()));
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(sizeof(float) * rows * cols));
}

std::string JsonEscape(const std::string& s) {
    std::string o;
    for (char c : s) {
        if (c == '"') o += "\\\"";
        else if (c == '\\') o += "\\\\";
        else if (c == '\n') o += "\\n";
        else o += c;
    }
    return o;
}

// ---------------------------------------------------------------------------- scenario
struct State {
    std::vector<double> x, y, z, vx, vy, vz, m;
    void Resize(int n) {
        for (auto* v : {&x, &y, &z, &vx, &vy, &vz, &m}) v->assign(static_cast<size_t>(n), 0.0);
    }
};

struct Scenario {
    State s;
    double G = 1.0, eps = 0.05, dt = 1e-3;
};

void Check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + " failed");
}

Scenario BuildScenario(const std::string& type, int n, uint32_t seed, double w0, double ecc, const std::string& icFile) {
    Scenario r;
    const double twoPi = 2.0 * M_PI;
    if (type == "plummer" || type == "plummer_sphere" || type == "hernquist" || type == "hernquist_sphere" ||
        type == "king" || type == "king_sphere") {
        n = std::max(n, 1);
        r.s.Resize(n);
        auto& s = r.s;
        const double spacing = 1.0 / std::cbrt(static_cast<double>(n));
        if (type.rfind("plummer", 0) == 0) {  // Scenarios.cpp:184-198
            Check(ngh_ic_plummer(n, seed, 1.0, 1.0, 1.0, s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                                 s.vy.data(), s.vz.data(), s.m.data()),
                  "ic::Plummer");
            r.eps = std::max(0.01, 1.5 * spacing);
        } else if (type.rfind("hernquist", 0) == 0) {  // Scenarios.cpp:219-237
            Check(ngh_ic_hernquist(n, seed, 1.0, 1.0, 1.0, s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                                   s.vy.data(), s.vz.data(), s.m.data()),
                  "ic::Hernquist");
            r.eps = std::max(0.01, 1.5 * spacing);
        } else {  // Scenarios.cpp:239-258
            Check(ngh_ic_king(n, seed, 1.0, 1.0, w0, 1.0, s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                              s.vy.data(), s.vz.data(), s.m.data()),
                  "ic::King");
            r.eps = std::max(0.01, 1.2 * spacing);
        }
        r.dt = twoPi / 800.0;
    } else if (type == "cluster" || type == "cold_collapse") {  // Scenarios.cpp:104-142, 201-214
        n = std::max(n, 1);
        r.s.Resize(n);
        auto& s = r.s;
        const bool cold = type == "cold_collapse";
        Check(ngh_ic_cluster(n, seed, cold ? 1 : 0, s.x.data(), s.y.data(), s.z.data(), s.vx.data(), s.vy.data(),
                             s.vz.data(), s.m.data()),
              "Cluster");
        const double spacing = 1.0 / std::cbrt(static_cast<double>(n));
        r.eps = cold ? std::max(0.01, 1.0 * spacing) : 2.0 * spacing;
        r.dt = cold ? 1.0 / 2000.0 : 1.0 / 1000.0;
    } else if (type == "two_body_kepler" || type == "kepler") {  // Scenarios.cpp:31-76
        const double G = 1.0, m = 1.0, sep = 1.0;
        const double vCirc = std::sqrt(G * 2.0 * m / sep);
        const double e = std::clamp(ecc, 0.0, 0.99);
        const double vRel = vCirc * std::sqrt(1.0 - e);
        r.s.Resize(2);
        r.s.x = {-sep / 2, sep / 2};
        r.s.vy = {-vRel / 2, vRel / 2};
        r.s.m = {m, m};
        r.eps = 0.02;
        const double a = sep / (1.0 + e);
        r.dt = twoPi * std::sqrt(a * a * a / (G * 2.0 * m)) / 2000.0;
    } else if (type == "lagrange_triangle" || type == "lagrange") {  // Scenarios.cpp:79-103
        const double G = 1.0, m = 1.0, L = 1.0;
        const double omega = std::sqrt(3.0 * G * m / (L * L * L)), R = L / std::sqrt(3.0);
        r.s.Resize(3);
        for (int k = 0; k < 3; ++k) {
            const double a = twoPi * k / 3.0;
            r.s.x[k] = R * std::cos(a);
            r.s.y[k] = R * std::sin(a);
            r.s.vx[k] = -omega * R * std::sin(a);
            r.s.vy[k] = omega * R * std::cos(a);
            r.s.m[k] = m;
        }
        r.eps = 0.02;
        r.dt = (twoPi / omega) / 2000.0;
    } else if (type == "ic_file") {  // Scenarios.cpp:287-329: N records of 7 float32 x,y,z,m,vx,vy,vz
        std::ifstream f(icFile, std::ios::binary | std::ios::ate);
        if (!f) throw std::runtime_error("ic_file: cannot open '" + icFile + "'");
        const std::streamsize bytes = f.tellg();
        f.seekg(0);
        const int cnt = static_cast<int>(bytes / (7 * static_cast<std::streamsize>(sizeof(float))));
        if (cnt <= 0) throw std::runtime_error("ic_file: '" + icFile + "' has no particles");
        std::vector<float> raw(static_cast<size_t>(cnt) * 7);
        f.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(raw.size() * sizeof(float)));
        r.s.Resize(cnt);
        for (int i = 0; i < cnt; ++i) {
            const float* p = &raw[static_cast<size_t>(i) * 7];
            r.s.x[i] = p[0];
            r.s.y[i] = p[1];
            r.s.z[i] = p[2];
            r.s.m[i] = p[3];
            r.s.vx[i] = p[4];
            r.s.vy[i] = p[5];
            r.s.vz[i] = p[6];
        }
        r.eps = 0.05;
        r.dt = 1e-3;
    } else {
        throw std::runtime_error("scenario '" + type + "' has no sampler in the GPU backend");
    }
    return r;
}

}  // namespace

extern "C" {

// Runs a deck on the GPU backend and writes the reference's output layout into out_dir.
// frames <= 0 takes time.frames from the deck (default 200).  Returns 0, or -1 with a
// message in err (capacity err_len) on failure.
int ngh_run_deck(const char* deck_path, const char* out_dir, int frames, int quiet, char* err, int err_len) {
    try {
        const Deck deck = Deck::FromFile(deck_path);
        // ---- DeckSim::Configure (DeckSim.cpp:26-74)
        const std::string title = deck.GetString("title", "N-Body Gravity");
        const std::string backend = deck.GetString("solver.backend", "gpu");
        if (backend != "gpu") throw std::runtime_error("solver.backend must be \"gpu\" for this runner");
        const std::string kind = deck.GetString("solver.kind", "barnes_hut");
        int solver;
        if (kind == "direct") solver = NBG_SOLVER_DIRECT;
        else if (kind == "barnes_hut" || kind == "bh") solver = NBG_SOLVER_BARNES_HUT;
        else throw std::runtime_error("solver.kind = \"" + kind + "\" is not supported by the GPU backend");
        const std::string mp = deck.GetString("solver.multipole", "quadrupole");
        if (mp != "quadrupole" && mp != "monopole") throw std::runtime_error("solver.multipole must be quadrupole|monopole");
        const int substeps = std::max(1, deck.GetInt("time.substeps_per_frame", 4));
        if (frames <= 0) frames = deck.GetInt("time.frames", 200);
        frames = std::max(frames, 1);

        Scenario sc = BuildScenario(deck.GetString("scenario.type", "plummer"), deck.GetInt("scenario.n", 2000),
                                    static_cast<uint32_t>(deck.GetInt("scenario.seed", 1)),
                                    deck.GetDouble("scenario.king_w0", 6.0), deck.GetDouble("scenario.eccentricity", 0.36),
                                    deck.GetString("scenario.ic_file", ""));
        // a scenario carries tuned G / softening / dt; the deck overrides each only when set
        // (DeckSim.cpp:53-60, 79-82)
        nbg_params p{};
        p.solver = solver;
        p.multipole = (mp == "quadrupole") ? NBG_QUADRUPOLE : NBG_MONOPOLE;
        p.G = deck.Has("sim.G") ? deck.GetDouble("sim.G", 1.0) : sc.G;
        const double eps = deck.Has("softening.eps") ? deck.GetDouble("softening.eps", 0.05) : sc.eps;
        p.eps2 = eps * eps;
        p.theta = deck.GetDouble("solver.theta", 0.5);
        p.dt = deck.Has("time.dt") ? deck.GetDouble("time.dt", 1e-3) : sc.dt;
        p.flags = NBG_FLAG_KAHAN;
        // softening kind, opening criterion and time-step scheme (DeckSim.cpp:30-41, 61-62, 81)
        p.softening = (deck.GetString("softening.kind", "plummer") == "spline") ? NBG_SOFT_SPLINE : NBG_SOFT_PLUMMER;
        p.mac = (deck.GetString("solver.mac", "geometric") == "relative") ? NBG_MAC_RELATIVE : NBG_MAC_GEOMETRIC;
        p.alpha = deck.GetDouble("solver.alpha", 0.001);
        p.adaptive = deck.GetBool("time.adaptive", false) ? 1 : 0;
        p.eta = deck.GetDouble("time.eta", 0.03);
        p.block = (deck.GetString("time.scheme", "global") == "block") ? 1 : 0;
        p.block_max_rung = deck.GetInt("time.block_max_rung", 8);

        const int n = static_cast<int>(sc.s.x.size());
        nbg_ctx* ctx = nbg_create(0, std::max(n, 1));
        if (!ctx) throw std::runtime_error("nbg_create failed");
        struct Closer {
            nbg_ctx* c;
            ~Closer() { nbg_destroy(c); }
        } closer{ctx};
        auto chk = [&](int rc, const char* what) {
            if (rc != NBG_OK) throw std::runtime_error(std::string(what) + ": " + nbg_last_error(ctx));
        };
        std::vector<float> f[7];
        const std::vector<double>* src[7] = {&sc.s.x, &sc.s.y, &sc.s.z, &sc.s.m, &sc.s.vx, &sc.s.vy, &sc.s.vz};
        for (int k = 0; k < 7; ++k) f[k].assign(src[k]->begin(), src[k]->end());
        chk(nbg_upload(ctx, n, f[0].data(), f[1].data(), f[2].data(), f[3].data(), f[4].data(), f[5].data(),
                       f[6].data()),
            "nbg_upload");
        double diag[11];
        chk(nbg_diagnostics(ctx, &p, 1, diag), "nbg_diagnostics");
        const double e0 = diag[0] + diag[1];
        const double l0 = std::sqrt(diag[5] * diag[5] + diag[6] * diag[6] + diag[7] * diag[7]);
        chk(nbg_accel(ctx, &p, nullptr), "nbg_accel");  // System::Prime

        // ---- fw::RunHeadless + OutputWriter
        const fs::path dir(out_dir);
        fs::create_directories(dir / "frames");
        {
            std::ofstream copy(dir / "deck.toml", std::ios::binary);
            copy << deck.Source();
        }
        const std::vector<std::string> names = {"energy", "energy_drift_pct", "L_mag", "L_drift_pct", "p_mag",
                                                "com_x",  "com_y",            "com_z", "dt_last",     "rung_max"};
        std::ofstream csv(dir / "diagnostics.csv", std::ios::binary);
        csv << "t,step";
        for (const auto& nm : names) csv << ',' << nm;
        csv << '\n';
        std::vector<float> x(n), y(n), z(n), vx(n), vy(n), vz(n), pos(3 * static_cast<size_t>(n)),
            vel(3 * static_cast<size_t>(n));
        char tag[16];
        double dt_last = p.dt;  // DeckSim::LoadScenario sets m_lastDt = dt (DeckSim.cpp:83)
        int rung_max = 0;
        for (int fr = 0; fr < frames; ++fr) {
            const long step = static_cast<long>(fr) * substeps;
            const double t = static_cast<double>(step) * p.dt;
            chk(nbg_download(ctx, x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(), nullptr, nullptr,
                             nullptr),
                "nbg_download");
            for (int i = 0; i < n; ++i) {
                pos[3 * i + 0] = x[i];
                pos[3 * i + 1] = y[i];
                pos[3 * i + 2] = z[i];
                vel[3 * i + 0] = vx[i];
                vel[3 * i + 1] = vy[i];
                vel[3 * i + 2] = vz[i];
            }
            std::snprintf(tag, sizeof(tag), "%04d", fr);
            WriteNpyF32(dir / "frames" / (std::string("pos_") + tag + ".npy"), pos.data(), n, 3);
            WriteNpyF32(dir / "frames" / (std::string("vel_") + tag + ".npy"), vel.data(), n, 3);
            chk(nbg_diagnostics(ctx, &p, 1, diag), "nbg_diagnostics");
            const double E = diag[0] + diag[1];
            const double L = std::sqrt(diag[5] * diag[5] + diag[6] * diag[6] + diag[7] * diag[7]);
            const double P = std::sqrt(diag[8] * diag[8] + diag[9] * diag[9] + diag[10] * diag[10]);
            const double row[10] = {E,
                                    (e0 != 0.0) ? 100.0 * (E - e0) / std::abs(e0) : 0.0,
                                    L,
                                    (l0 > 1e-12) ? 100.0 * (L - l0) / l0 : 100.0 * L,
                                    P,
                                    diag[2],
                                    diag[3],
                                    diag[4],
                                    dt_last,
                                    static_cast<double>(rung_max)};
            csv << std::setprecision(10) << t << ',' << step;
            for (double v : row) csv << ',' << v;
            csv << '\n';
            csv.flush();
            if (!quiet) {
                std::printf("\rframe %d/%d  t=%.4f   ", fr + 1, frames, t);
                std::fflush(stdout);
            }
            if (fr + 1 < frames) {
                nbg_stats st{};
                chk(nbg_step(ctx, &p, substeps, &st), "nbg_step");
                dt_last = st.dt_last;
                rung_max = st.rung_max;
            }
        }
        csv.close();
        std::ofstream j(dir / "manifest.json", std::ios::binary);
        j << "{\n  \"title\": \"" << JsonEscape(title) << "\",\n";
        j << "  \"grid\": { \"nx\": 0, \"ny\": 0, \"lx\": 0, \"ly\": 0 },\n";
        j << "  \"dt\": " << p.dt << ",\n  \"substeps_per_frame\": " << substeps << ",\n";
        j << "  \"frames\": " << frames << ",\n  \"deck_file\": \"deck.toml\",\n";
        j << "  \"fields\": {\n    \"pos\": { \"dtype\": \"<f4\", \"shape\": [" << n << ", 3] },\n";
        j << "    \"vel\": { \"dtype\": \"<f4\", \"shape\": [" << n << ", 3] }\n  },\n";
        j << "  \"diagnostics\": [";
        for (size_t i = 0; i < names.size(); ++i) j << (i ? ", " : "") << '"' << names[i] << '"';
        j << "]\n}\n";
        if (!quiet) std::printf("\ndone -> %s\n", out_dir);
        return 0;
    } catch (const std::exception& e) {
        if (err && err_len > 0) std::snprintf(err, static_cast<size_t>(err_len), "%s", e.what());
        return -1;
    }
}

}  // extern "C"
