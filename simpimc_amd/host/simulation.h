// simulation.h -- host side of the engine: simpimc's Path / Action / Move /
// Observable / Loop plug-ins for the hot path, forwarding to the C ABI.
//
// The classes keep the reference's names, XML attributes, defaults and call
// structure (framework/simulation_class.h:16-45, events/loop_class.h:34-60):
// actions are built from <Actions>, moves from <Moves>, observables from
// <Observables>, and the <Algorithm> loop runs them.  Compute goes to the
// device through include/spimc.h; a Loop whose only event is one Move hands
// its n_step attempts to the device as one stage.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spimc.h"
#include "table_file.h"
#include "xml_input.h"

namespace spimc_host {

inline void check(int rc, const char* what) {
    if (rc != SPIMC_OK)
        throw std::runtime_error(std::string(what) + ": " + spimc_last_error() + " (code " + std::to_string(rc) + ")");
}

// ---------------------------------------------------------------------------
// Path: <System> and one <Species> (data_structures/path_class.h:25-69,
// species_class.h:48-98).
// ---------------------------------------------------------------------------
struct Path {
    uint32_t n_d = 3, n_bead = 0, n_part = 0;
    double beta = 0., L = 0., tau = 0., lambda = 0., k_cut = 0.;
    bool pbc = true;
    std::string species;
    Input species_in;
    std::vector<int32_t> k_n;  // the k-vector set, grown by SetupK

    explicit Path(const Input& in) {
        const Input& sys = in.GetChild("System");
        n_d = sys.GetAttribute<uint32_t>("n_d");
        n_bead = sys.GetAttribute<uint32_t>("n_bead");
        beta = sys.GetAttribute<double>("beta");
        pbc = sys.GetAttribute<bool>("pbc", true);
        L = sys.GetAttribute<double>("L", 0.);
        tau = beta / (1. * n_bead);
        auto sp = in.GetChild("Particles").GetChildList("Species");
        if (sp.size() != 1) throw std::runtime_error("exactly one <Species> is supported");
        species_in = sp[0];
        species = sp[0].GetAttribute<std::string>("name");
        n_part = sp[0].GetAttribute<uint32_t>("n_part");
        lambda = sp[0].GetAttribute<double>("lambda");
        if (sp[0].GetAttribute<bool>("fermi", false)) throw std::runtime_error("fermions are outside this engine");
        if (!pbc || !(L > 0.)) throw std::runtime_error("a periodic box (pbc=1, L>0) is required");
        SetupK(sys.GetAttribute<double>("k_cut", 2. * M_PI / L));
    }

    // KSpace::Setup grows the set only (k_space_class.h:33-40)
    void SetupK(double kc) {
        if (kc <= k_cut) return;
        k_cut = kc;
        uint32_t n = 0;
        check(spimc_kspace_build(L, kc, nullptr, 0, &n), "kspace");
        k_n.assign(3 * (size_t)n, 0);
        check(spimc_kspace_build(L, kc, k_n.data(), n, &n), "kspace");
    }

    uint32_t n_k() const { return (uint32_t)(k_n.size() / 3); }
    double kmag(uint32_t k) const {
        const double b = 2. * M_PI / L;
        double x = k_n[3 * k] * b, y = k_n[3 * k + 1] * b, z = k_n[3 * k + 2] * b;
        return std::sqrt(x * x + y * y + z * z);
    }

    // Initial path of one replica, positions r[3][M][N] (species_class.h:222-377)
    // sigma receives the path permutation (the identity except for a restart)
    std::vector<double> InitPath(uint64_t seed, uint32_t replica, std::vector<uint32_t>& sigma) const {
        sigma.resize(n_part);
        for (uint32_t p = 0; p < n_part; ++p) sigma[p] = p;
        const std::string init = species_in.GetAttribute<std::string>("init_type");
        std::vector<double> r((size_t)3 * n_bead * n_part);
        auto set = [&](uint32_t p, const double* x) {
            for (uint32_t d = 0; d < 3; ++d)
                for (uint32_t b = 0; b < n_bead; ++b) r[((size_t)d * n_bead + b) * n_part + p] = x[d];
        };
        if (init == "Random") {
            const double cof = species_in.GetAttribute<double>("cofactor", 1.);
            std::mt19937_64 g(seed);
            std::uniform_real_distribution<double> u(-cof * L / 2., cof * L / 2.);
            for (uint32_t p = 0; p < n_part; ++p) {
                double x[3] = {u(g), u(g), u(g)};
                set(p, x);
            }
        } else if (init == "BCC" || init == "Cubic") {
            const int n = (int)std::ceil(std::pow(0.5 * n_part, 1.0 / 3));
            const double cof = init == "Cubic" ? species_in.GetAttribute<double>("cofactor", 1.) : 1.;
            const double delta = cof * L / n;
            for (uint32_t p_i = 0; p_i < n_part; ++p_i) {
                const int p = init == "BCC" ? (int)p_i / 2 : (int)p_i;
                int t0 = p / (n * n), t1 = (p - t0 * n * n) / n, t2 = p - t0 * n * n - t1 * n;
                double x[3] = {delta * t0 - 0.5 * L, delta * t1 - 0.5 * L, delta * t2 - 0.5 * L};
                if (init == "BCC" && (p_i % 2))
                    for (auto& v : x) v += 0.5 * delta;
                set(p_i, x);
            }
        } else if (init == "Line") {
            const double dr = species_in.GetAttribute<double>("spacing", L / n_part);
            for (uint32_t p = 0; p < n_part; ++p) {
                double x[3] = {p * dr, 0., 0.};
                set(p, x);
            }
        } else if (init == "File") {
            const std::string fn = species_in.GetAttribute<std::string>("file_name");
            const bool all = species_in.GetAttribute<bool>("all_beads", false);
            std::ifstream f(fn);
            if (!f) throw std::runtime_error("Init file '" + fn + "' does not exist.");
            for (uint32_t p = 0; p < n_part; ++p) {
                if (all) {
                    for (uint32_t b = 0; b < n_bead; ++b)
                        for (uint32_t d = 0; d < 3; ++d) f >> r[((size_t)d * n_bead + b) * n_part + p];
                } else {
                    double x[3];
                    f >> x[0] >> x[1] >> x[2];
                    set(p, x);
                }
            }
        } else if (init == "Restart") {
            RestartPath(r, sigma, replica);
        } else {
            throw std::runtime_error("init_type " + init + " is outside this engine");
        }
        return r;
    }

    // init_type "Restart" (species_class.h:336-377): the last dump of a PathDump
    // observable in <prefix>.<replica>.h5 (parallel="1") or <prefix>.0.h5.
    // Positions are read as cube(n_d, n_part * n_bead, n_dump), entry
    // (d, p * n_bead + b, n_dump - 1); the permutation of the last dump gives, per
    // particle, the previous particle of its first bead and the next particle of
    // its last bead (species_class.h:366-374).
    void RestartPath(std::vector<double>& r, std::vector<uint32_t>& sigma, uint32_t replica) const {
        const std::string prefix = species_in.GetAttribute<std::string>("prefix");
        const std::string dump = species_in.GetAttribute<std::string>("path_dump_name", "path_dump");
        const bool parallel = species_in.GetAttribute<bool>("parallel", false);
        const std::string file = prefix + "." + std::to_string(parallel ? replica : 0) + ".h5";
        const TableData t = ReadHdf5(file);
        const std::string g = "Observables/" + dump + "/" + species + "/";
        auto get = [&](const std::string& n) -> const std::vector<double>& {
            auto it = t.data.find(g + n);
            if (it == t.data.end()) throw std::runtime_error(file + ": missing dataset " + g + n);
            return it->second;
        };
        const auto& nd = get("n_dump");
        if (nd.empty() || nd[0] < 1) throw std::runtime_error(file + ": no path dump");
        const size_t n_dump = (size_t)nd[0];
        const auto& pos = get("positions");
        const auto& perm = get("permutation");
        const size_t per = (size_t)n_d * n_part * n_bead;
        if (pos.size() < per * n_dump || perm.size() < (size_t)2 * n_part * n_dump)
            throw std::runtime_error(file + ": path dump smaller than n_dump paths of this system");
        const double* last = pos.data() + per * (n_dump - 1);
        for (uint32_t p = 0; p < n_part; ++p)
            for (uint32_t b = 0; b < n_bead; ++b)
                for (uint32_t d = 0; d < n_d; ++d)
                    r[((size_t)d * n_bead + b) * n_part + p] = last[d + (size_t)n_d * (p * (size_t)n_bead + b)];
        const double* pl = perm.data() + (size_t)2 * n_part * (n_dump - 1);
        for (uint32_t p = 0; p < n_part; ++p) {
            const double nx = pl[2 * p + 1];
            if (!(nx >= 0. && nx < n_part)) throw std::runtime_error(file + ": permutation out of range");
            sigma[p] = (uint32_t)nx;
        }
        for (uint32_t p = 0; p < n_part; ++p)
            if (pl[2 * sigma[p]] != p) throw std::runtime_error(file + ": permutation rows disagree");
        std::cout << "Restarting " << species << " from " << file << "..." << std::endl;
    }
};

// ---------------------------------------------------------------------------
// Actions (actions/action_class.h, actions/actions.h:13-35)
// ---------------------------------------------------------------------------
struct Action {
    std::string name, type;
    int n_images = 0;
    uint32_t max_level = 0;
    explicit Action(const Input& in)
        : name(in.GetAttribute<std::string>("name")),
          type(in.GetAttribute<std::string>("type")),
          n_images(in.GetAttribute<int>("n_images", 0)),
          max_level(in.GetAttribute<uint32_t>("max_level", 0)) {}
    virtual ~Action() = default;
    virtual void Fill(spimc_system_desc& d, Path& path) = 0;
    virtual int Selector() const = 0;
};

struct Kinetic : Action {
    explicit Kinetic(const Input& in) : Action(in) {}
    void Fill(spimc_system_desc& d, Path&) override { d.kinetic_n_images = n_images; }
    int Selector() const override { return SPIMC_ACTION_KINETIC; }
};

struct PairActionBase : Action {
    bool use_long_range = false;
    uint32_t n_order = 0;
    std::string file, species_a, species_b;
    double k_cut = 0.;
    TableFile table;
    std::vector<double> v_long_k;
    explicit PairActionBase(const Input& in, Path& path) : Action(in) {
        n_order = in.GetAttribute<uint32_t>("n_order", 0);
        species_a = in.GetAttribute<std::string>("species_a");
        species_b = in.GetAttribute<std::string>("species_b");
        if (species_a != path.species || species_b != path.species)
            throw std::runtime_error("pair action " + name + ": only the homologous pair of the one species is supported");
        use_long_range = in.GetAttribute<bool>("use_long_range", false);
        if (use_long_range) {
            k_cut = in.GetAttribute<double>("k_cut", path.k_cut);
            path.SetupK(k_cut);
        }
        if (max_level != 0) throw std::runtime_error("pair action " + name + ": max_level must be 0");
        file = in.GetAttribute<std::string>("file");
        table.Load(file);
    }
    int Selector() const override { return SPIMC_ACTION_PAIR; }
    // per stored k: the table entry whose |k| equals it within 1e-8 (bare_pair_action_class.h:79-85)
    void MatchK(const Path& path, const std::vector<double>& k_v, const std::vector<double>& vals, double scale) {
        v_long_k.assign(path.n_k(), 0.);
        for (uint32_t k = 0; k < path.n_k(); ++k)
            for (size_t i = 0; i < k_v.size(); ++i)
                if (std::fabs(path.kmag(k) - k_v[i]) < 1e-8) v_long_k[k] = vals[i] * scale;
    }
};

// BarePairAction loader (bare_pair_action_class.h:22-96)
struct BarePairAction : PairActionBase {
    std::vector<double> r, v_r, r_long, v_long_r;
    double v_long_k_0 = 0., v_long_r_0 = 0.;
    bool is_coulomb;
    BarePairAction(const Input& in, Path& path) : PairActionBase(in, path) {
        is_coulomb = in.GetAttribute<bool>("is_coulomb", false);
        r = table.Read("v/diag/r");
        v_r = table.Read("v/diag/v_r");
        if (use_long_range) {
            r_long = table.Read("v/diag/r_long");
            v_long_r = table.Read("v/diag/v_long_r");
            v_long_r_0 = table.Scalar("v/diag/v_long_r_0");
            v_long_k_0 = table.Scalar("v/diag/v_long_k_0");
        }
    }
    void Fill(spimc_system_desc& d, Path& path) override {
        d.pair_type = SPIMC_PAIR_BARE;
        d.pair_use_long_range = use_long_range;
        d.bare_is_coulomb = is_coulomb;
        d.bare_n_r = (uint32_t)r.size();
        d.bare_r = r.data();
        d.bare_v_r = v_r.data();
        if (use_long_range) {
            MatchK(path, table.Read("v/diag/k"), table.Read("v/diag/v_long_k"), 1.);
            d.bare_n_r_long = (uint32_t)r_long.size();
            d.bare_r_long = r_long.data();
            d.bare_v_long_r = v_long_r.data();
            d.pair_v_long_k = v_long_k.data();
            d.pair_v_long_k_0 = v_long_k_0;
            d.pair_v_long_r_0 = v_long_r_0;
        }
    }
};

// DavidPairAction loader (david_pair_action_class.h:192-360)
struct DavidPairAction : PairActionBase {
    std::vector<double> grid, u_kj, du_kj;
    double v_long_k_0 = 0., v_long_r_0 = 0.;
    DavidPairAction(const Input& in, Path& path) : PairActionBase(in, path) {
        const std::string pre = "u_kj_" + std::to_string(n_order), dpre = "du_kj_dbeta_" + std::to_string(n_order);
        const double start = table.Scalar(pre + "/grid/start"), end = table.Scalar(pre + "/grid/end");
        const uint32_t n_grid = (uint32_t)table.Scalar(pre + "/grid/n_grid_points");
        const std::string gtype = table.String(pre + "/grid/type");
        if (gtype.find("LINEAR") != std::string::npos) {
            grid.resize(n_grid);
            for (uint32_t i = 0; i < n_grid; ++i) grid[i] = start + i * (end - start) / (n_grid - 1);
        } else if (table.Has(pre + "/grid/grid_points")) {
            grid = table.Read(pre + "/grid/grid_points");
        } else {
            throw std::runtime_error("grid type " + gtype + " needs grid_points in this engine");
        }
        // tau must be one of the table's taus within 1e-6 (david_pair_action_class.h:236-245)
        const auto& taus = table.Read(pre + "/taus");
        const uint32_t n_tau = max_level + 1;
        int tau_i = -1;
        for (uint32_t t = 0; t < n_tau && t < taus.size(); ++t)
            if (std::fabs(taus[t] - path.tau) < 1.0e-6) tau_i = (int)t;
        if (tau_i < 0) throw std::runtime_error("tau of " + std::to_string(path.tau) + " not found.");
        uint32_t n_val = 1;
        for (uint32_t i = 1; i <= n_order; ++i) n_val += 1 + i;
        const auto& V = table.Read("potential/data");
        // the cubes span every tau the file holds; tau is their slowest index
        const std::vector<double> u = table.Cube(pre + "/data", n_val, n_grid, taus.size());
        const std::vector<double> du = table.Cube(dpre + "/data", n_val, n_grid, taus.size());
        if (V.size() < n_grid) throw std::runtime_error("potential/data holds fewer than n_grid_points values");
        // component 0 = potential; knots 0 .. n_grid-2 filled, the last one left 0
        u_kj.assign((size_t)(n_val + 1) * n_grid, 0.);
        du_kj.assign((size_t)(n_val + 1) * n_grid, 0.);
        for (uint32_t g = 0; g + 1 < n_grid; ++g) {
            u_kj[g] = V[g];
            du_kj[g] = V[g];
            for (uint32_t c = 1; c <= n_val; ++c) {
                const size_t src = (c - 1) + (size_t)n_val * (g + (size_t)n_grid * tau_i);
                u_kj[(size_t)c * n_grid + g] = u[src];
                du_kj[(size_t)c * n_grid + g] = du[src];
            }
        }
        if (use_long_range) {
            v_long_r_0 = table.Scalar("squarer/v_image");
            const auto& kp = table.Read("long_range/k_points");
            const auto& uk = table.Read("long_range/u_k");
            const double vol = path.L * path.L * path.L;
            for (size_t i = 0; i < kp.size(); ++i)
                if (std::fabs(kp[i]) < 1e-8) v_long_k_0 = uk[i] / vol;
        }
    }
    void Fill(spimc_system_desc& d, Path& path) override {
        d.pair_type = SPIMC_PAIR_DAVID;
        d.pair_use_long_range = use_long_range;
        d.david_n_order = n_order;
        d.david_n_grid = (uint32_t)grid.size();
        d.david_grid = grid.data();
        d.david_u_kj = u_kj.data();
        d.david_du_kj = du_kj.data();
        if (use_long_range) {
            const double vol = path.L * path.L * path.L;
            MatchK(path, table.Read("long_range/k_points"), table.Read("long_range/u_k"), 1. / vol);
            d.pair_v_long_k = v_long_k.data();
            d.pair_v_long_k_0 = v_long_k_0;
            d.pair_v_long_r_0 = v_long_r_0;
        }
    }
};

// IlkkaPairAction loader (ilkka_pair_action_class.h:252-431)
struct IlkkaPairAction : PairActionBase {
    std::vector<double> x_u, y_u, u_xy, x_du, y_du, du_xy, r_v, v_r;
    std::vector<double> r_v_long, v_long_r, r_u_long, u_long_r, r_du_long, du_long_r;
    std::vector<double> u_long_k, du_long_k;
    double v_long_k_0 = 0., v_long_r_0 = 0., du_long_k_0 = 0., du_long_r_0 = 0.;
    IlkkaPairAction(const Input& in, Path& path) : PairActionBase(in, path) {
        auto grid2 = [&](const std::string& g, const std::string& v, std::vector<double>& x, std::vector<double>& y,
                         std::vector<double>& xy) {
            const size_t n_x = (size_t)table.Scalar(g + "/off_diag/n_x"), n_y = (size_t)table.Scalar(g + "/off_diag/n_y");
            x = table.Read(g + "/off_diag/x");
            y = table.Read(g + "/off_diag/y");
            if (x.size() != n_x || y.size() != n_y) throw std::runtime_error(g + "/off_diag: grid sizes differ from n_x, n_y");
            xy = table.Matrix(g + "/off_diag/" + v, n_x, n_y);
        };
        grid2("u", "u_xy", x_u, y_u, u_xy);
        grid2("du", "du_xy", x_du, y_du, du_xy);
        r_v = table.Read("v/diag/r");
        v_r = table.Read("v/diag/v_r");
        if (use_long_range) {
            r_u_long = table.Read("u/diag/r_long");
            u_long_r = table.Read("u/diag/u_long_r");
            r_du_long = table.Read("du/diag/r_long");
            du_long_r = table.Read("du/diag/du_long_r");
            du_long_r_0 = table.Scalar("du/diag/du_long_r_0");
            du_long_k_0 = table.Scalar("du/diag/du_long_k_0");
            r_v_long = table.Read("v/diag/r_long");
            v_long_r = table.Read("v/diag/v_long_r");
            v_long_r_0 = table.Scalar("v/diag/v_long_r_0");
            v_long_k_0 = table.Scalar("v/diag/v_long_k_0");
        }
    }
    void Fill(spimc_system_desc& d, Path& path) override {
        d.pair_type = SPIMC_PAIR_ILKKA;
        d.pair_use_long_range = use_long_range;
        d.ilkka_n_x_u = (uint32_t)x_u.size();
        d.ilkka_n_y_u = (uint32_t)y_u.size();
        d.ilkka_x_u = x_u.data();
        d.ilkka_y_u = y_u.data();
        d.ilkka_u_xy = u_xy.data();
        d.ilkka_n_x_du = (uint32_t)x_du.size();
        d.ilkka_n_y_du = (uint32_t)y_du.size();
        d.ilkka_x_du = x_du.data();
        d.ilkka_y_du = y_du.data();
        d.ilkka_du_xy = du_xy.data();
        d.ilkka_n_r_v = (uint32_t)r_v.size();
        d.ilkka_r_v = r_v.data();
        d.ilkka_v_r = v_r.data();
        if (use_long_range) {
            d.ilkka_n_r_v_long = (uint32_t)r_v_long.size();
            d.ilkka_r_v_long = r_v_long.data();
            d.ilkka_v_long_r = v_long_r.data();
            d.ilkka_n_r_u_long = (uint32_t)r_u_long.size();
            d.ilkka_r_u_long = r_u_long.data();
            d.ilkka_u_long_r = u_long_r.data();
            d.ilkka_n_r_du_long = (uint32_t)r_du_long.size();
            d.ilkka_r_du_long = r_du_long.data();
            d.ilkka_du_long_r = du_long_r.data();
            // per stored k: the table entry whose |k| equals it within 1e-8 (ilkka_pair_action_class.h:301-308)
            MatchK(path, table.Read("u/diag/k"), table.Read("u/diag/u_long_k"), 1.);
            u_long_k = v_long_k;
            MatchK(path, table.Read("du/diag/k"), table.Read("du/diag/du_long_k"), 1.);
            du_long_k = v_long_k;
            MatchK(path, table.Read("v/diag/k"), table.Read("v/diag/v_long_k"), 1.);
            d.ilkka_u_long_k = u_long_k.data();
            d.ilkka_du_long_k = du_long_k.data();
            d.pair_v_long_k = v_long_k.data();
            d.pair_v_long_k_0 = v_long_k_0;
            d.pair_v_long_r_0 = v_long_r_0;
            d.ilkka_du_long_k_0 = du_long_k_0;
            d.ilkka_du_long_r_0 = du_long_r_0;
        }
    }
};

// ---------------------------------------------------------------------------
// Events (events/event_class.h, events/moves/move_class.h, observables/*)
// ---------------------------------------------------------------------------
struct Engine {
    spimc_ctx* ctx = nullptr;
    uint32_t n_replicas = 1;
    const void* perm_owner = nullptr;  // the permuting move whose level and epsilon the engine holds
    ~Engine() {
        if (ctx) spimc_ctx_destroy(ctx);
    }
};

struct Event {
    std::string name;
    double time_spent = 0.;
    virtual ~Event() = default;
    virtual void DoEvent() = 0;
    // A run of n consecutive DoEvent calls; moves hand them to the device at once.
    virtual void DoEvents(uint32_t n) {
        for (uint32_t i = 0; i < n; ++i) DoEvent();
    }
    double Now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }
};

struct Move : Event {
    Engine& eng;
    std::string type;
    std::vector<uint64_t> last;  // counters at the last Write
    Move(const Input& in, Engine& e) : eng(e), type(in.GetAttribute<std::string>("type")) {
        name = in.GetAttribute<std::string>("name");
    }
    virtual ~Move() = default;
    virtual int Stage(uint32_t n) = 0;
    virtual int CounterBase() const = 0;
    virtual void Counters(uint32_t r, uint64_t* c) const { check(spimc_move_counters(eng.ctx, r, c), "counters"); }
    void DoEvent() override { DoEvents(1); }
    void DoEvents(uint32_t n) override {
        double t = Now();
        check(Stage(n), name.c_str());
        time_spent += Now() - t;
    }
    // Move::Write (move_class.h:80-94): per replica, attempts and acceptances
    // since the last write
    struct Counts {
        uint64_t n_attempt, n_accept;
    };
    std::vector<Counts> Block() {
        std::vector<Counts> out(eng.n_replicas);
        last.resize(2 * eng.n_replicas, 0);
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            uint64_t c[4];
            Counters(r, c);
            out[r] = {c[CounterBase()] - last[2 * r], c[CounterBase() + 1] - last[2 * r + 1]};
            last[2 * r] = c[CounterBase()];
            last[2 * r + 1] = c[CounterBase() + 1];
        }
        return out;
    }
};

// Bisect (bisect_class.h:167-190)
struct Bisect : Move {
    bool adaptive = false;
    double target_ratio = 0.;
    uint32_t n_bead = 0, n_level0 = 0;
    std::vector<uint32_t> n_level;  // per replica, the level of the running block
    bool first_write = true;
    int n_images = 0;
    Bisect(const Input& in, Engine& e, const Path& path) : Move(in, e) {
        n_level0 = in.GetAttribute<uint32_t>("n_level");
        n_images = in.GetAttribute<int>("n_images", 0);
        adaptive = in.GetAttribute<bool>("adaptive", false);
        if (adaptive) target_ratio = in.GetAttribute<double>("target_ratio");
        n_bead = path.n_bead;
    }
    int Stage(uint32_t n) override { return spimc_bisect_stage(eng.ctx, nullptr, n, nullptr); }
    int CounterBase() const override { return 0; }
    virtual void SetLevel(uint32_t r, uint32_t lv) { check(spimc_set_bisect_level(eng.ctx, r, lv), "adaptive bisection level"); }
    // Bisect::Reset with adaptive="1" (bisect_class.h:143-151), after the first
    // write: one level down when the block's acceptance is below target_ratio,
    // else one up while 2^n_level < n_bead / 2; each replica on its own counts
    void Adapt(const std::vector<Counts>& b) {
        n_level.resize(eng.n_replicas, n_level0);
        if (!adaptive || first_write) {
            first_write = false;
            return;
        }
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            const double ratio = (double)b[r].n_accept / (double)b[r].n_attempt;
            uint32_t lv = n_level[r];
            if (ratio < target_ratio && lv > 1)
                --lv;
            else if ((1u << lv) < n_bead / 2)
                ++lv;
            if (lv != n_level[r]) SetLevel(r, lv);
            n_level[r] = lv;
        }
    }
};

// PermBisectIterative (perm_bisect_iterative_class.h:226-230, perm_bisect_class.h:100-116):
// the bisection attributes plus epsilon; its own level per replica, and the
// per-cycle-length counters of each block (perm_bisect_class.h:119-132).  The
// engine keeps one set of permutation counters per replica; each move adds to
// its own copy what its stages changed, so two permuting moves keep apart.
struct PermBisectIterative : Bisect {
    double epsilon;
    uint32_t n_part;
    std::vector<uint64_t> own;                          // [replica][2 + 2 n_part] this move's counts
    std::vector<uint64_t> perm_last;                    // [replica][2 + 2 n_part] at the last write
    std::vector<std::vector<uint64_t>> perm_att, perm_acc;  // the block's counts per replica
    bool h5_started = false;
    PermBisectIterative(const Input& in, Engine& e, const Path& path) : Bisect(in, e, path) {
        epsilon = in.GetAttribute<double>("epsilon", 1.e-100);
        n_part = path.n_part;
    }
    virtual int PermStage(uint32_t n) { return spimc_perm_bisect_stage(eng.ctx, nullptr, n, nullptr); }
    // PermBisect::n_perm_type: N for the iterative move (perm_bisect_iterative_class.h:229)
    virtual int NPermType() const { return (int)n_part; }
    std::vector<uint64_t> ReadAll() const {
        const size_t w = 2 + 2 * (size_t)n_part;
        std::vector<uint64_t> v(eng.n_replicas * w);
        for (uint32_t r = 0; r < eng.n_replicas; ++r) check(spimc_perm_counters(eng.ctx, r, &v[r * w]), "counters");
        return v;
    }
    // the engine holds one permuting level and epsilon per replica: a move takes
    // them over (its own, adapted level per replica) before its stage
    void Activate() {
        for (uint32_t r = 0; r < eng.n_replicas; ++r)
            check(spimc_perm_configure(eng.ctx, r, n_level.empty() ? n_level0 : n_level[r], epsilon), name.c_str());
        eng.perm_owner = this;
    }
    int Stage(uint32_t n) override {
        if (eng.perm_owner != this) Activate();
        const auto before = ReadAll();
        const int rc = PermStage(n);
        if (rc) return rc;
        const auto after = ReadAll();
        own.resize(after.size(), 0);
        for (size_t i = 0; i < after.size(); ++i) own[i] += after[i] - before[i];
        return rc;
    }
    void Counters(uint32_t r, uint64_t* c) const override {
        const size_t w = 2 + 2 * (size_t)n_part;
        c[0] = own.empty() ? 0 : own[r * w];
        c[1] = own.empty() ? 0 : own[r * w + 1];
    }
    void SetLevel(uint32_t r, uint32_t lv) override {
        if (eng.perm_owner == this) check(spimc_perm_configure(eng.ctx, r, lv, epsilon), "adaptive bisection level");
    }
    virtual void Configure() { Activate(); }
    // perm_attempt / perm_accept since the last write (PermBisect::Reset zeroes them)
    void PermBlock() {
        const size_t w = 2 + 2 * (size_t)n_part;
        perm_last.resize(eng.n_replicas * w, 0);
        own.resize(eng.n_replicas * w, 0);
        perm_att.assign(eng.n_replicas, std::vector<uint64_t>(n_part));
        perm_acc.assign(eng.n_replicas, std::vector<uint64_t>(n_part));
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            const uint64_t* v = &own[r * w];
            uint64_t* l = &perm_last[r * w];
            for (uint32_t c = 0; c < n_part; ++c) {
                perm_att[r][c] = v[2 + c] - l[2 + c];
                perm_acc[r][c] = v[2 + n_part + c] - l[2 + n_part + c];
            }
            std::copy(v, v + w, l);
        }
    }
};

// PermBisectTable (perm_bisect_table_class.h:231-239): PermBisect's attributes
// plus n_perm_part.  detailed_balance="1" (no reference counterpart) selects the
// final table check perm_tot_0 / (perm_tot_1 w) in place of the reference's
// perm_tot_0 / perm_tot_1, which is biased when n_perm_part < n_part.
struct PermBisectTable : PermBisectIterative {
    uint32_t n_perm_part;
    uint32_t balance;
    PermBisectTable(const Input& in, Engine& e, const Path& path) : PermBisectIterative(in, e, path) {
        n_perm_part = in.GetAttribute<uint32_t>("n_perm_part");
        balance = in.GetAttribute<bool>("detailed_balance", false) ? 1u : 0u;
    }
    int PermStage(uint32_t n) override { return spimc_perm_table_stage(eng.ctx, nullptr, n, nullptr); }
    // the permutation types of n_perm_part particles (BuildCycles, perm_bisect_table_class.h:61)
    int NPermType() const override {
        int f = 1;
        for (uint32_t i = 2; i <= n_perm_part; ++i) f *= (int)i;
        return f;
    }
    void Configure() override {
        PermBisectIterative::Configure();
        for (uint32_t r = 0; r < eng.n_replicas; ++r)
            check(spimc_perm_table_configure(eng.ctx, r, n_perm_part, balance), name.c_str());
    }
};

// DisplaceParticle (displace_particle_class.h:93)
struct DisplaceParticle : Move {
    DisplaceParticle(const Input& in, Engine& e, spimc_system_desc& d, const Path& path) : Move(in, e) {
        d.displace_step = in.GetAttribute<double>("step_size", path.L / 10.);
    }
    int Stage(uint32_t n) override { return spimc_displace_stage(eng.ctx, nullptr, n, nullptr); }
    int CounterBase() const override { return 2; }
};

struct Observable : Event {
    Engine& eng;
    std::string type;
    uint32_t which = 0;
    Observable(const Input& in, Engine& e) : eng(e), type(in.GetAttribute<std::string>("type")) {
        name = in.GetAttribute<std::string>("name");
    }
    void DoEvent() override {
        double t = Now();
        check(spimc_obs_accumulate(eng.ctx, which), name.c_str());
        time_spent += Now() - t;
    }
};

// Energy (energy_class.h:179-230, Write :232-325 without sectors)
struct Energy : Observable {
    std::vector<std::string> action_names;
    int estimator;
    Energy(const Input& in, Engine& e, const std::vector<std::string>& names) : Observable(in, e), action_names(names) {
        which = SPIMC_OBS_ENERGY;
        const bool virial = in.GetAttribute<bool>("use_virial_estimator", false);
        // use_ewald_virial selects the analytic adjoint virial (no reference counterpart)
        estimator = virial ? (in.GetAttribute<bool>("use_ewald_virial", false) ? SPIMC_ENERGY_VIRIAL_ADJOINT
                                                                                : SPIMC_ENERGY_VIRIAL_COMPAT)
                           : SPIMC_ENERGY_THERMAL;
        if (in.GetAttribute<bool>("measure_per_sector", false) || in.GetAttribute<bool>("measure_potential", false))
            throw std::runtime_error("Energy: measure_per_sector/measure_potential are outside this engine");
    }
    // per replica: the kinetic and pair energies per bead and measurement
    // since the last write, and the number of measurements; resets the sums
    struct Energies {
        double kinetic, pair;
        uint64_t n_measure;
    };
    std::vector<Energies> Block(uint32_t n_bead) {
        std::vector<Energies> out(eng.n_replicas);
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            double e[2];
            uint64_t nm = 0;
            check(spimc_obs_read(eng.ctx, r, SPIMC_OBS_ENERGY, e, &nm), "energy");
            const double norm = (double)n_bead * (nm ? nm : 1);
            out[r] = {e[0] / norm, e[1] / norm, nm};
        }
        check(spimc_obs_reset(eng.ctx, SPIMC_OBS_ENERGY), "reset");
        return out;
    }
};

struct StructureFactor : Observable {
    double k_cut;
    std::string species_a, species_b;
    StructureFactor(const Input& in, Engine& e, Path& path) : Observable(in, e) {
        which = SPIMC_OBS_SK;
        species_a = in.GetAttribute<std::string>("species_a", path.species);
        species_b = in.GetAttribute<std::string>("species_b", path.species);
        k_cut = in.GetAttribute<double>("k_cut", path.k_cut);
        path.SetupK(k_cut);
    }
};

// Time (time_class.h:7-71): at every Write, the block's wall time and each
// event's time spent in it, then a fresh block.  Moves time the host call that
// hands their attempts to the device, which returns when the stage is done.
struct TimeObs : Observable {
    double start;
    TimeObs(const Input& in, Engine& e) : Observable(in, e) { start = Now(); }
    void DoEvent() override {}
};

// Permutation (permutation_class.h:7-119): per measurement, the cycle lengths
// of each replica's path permutation (Species::SetCycleCount,
// species_class.h:28-44) and its sector, the index of the sorted cycle lengths
// among the integer partitions of n_part in the order of
// Species::SetupPermSectors (species_class.h:178-219, ascending compositions).
// The reference never fills its partition map (the setup runs only when the
// map is already non-empty), so its sector lookup reads an empty map; here the
// map is built as the setup describes, for n_part <= 60 (more partitions than
// that are not enumerated: the sector series is then left out).
struct Permutation : Observable {
    std::string species;
    uint32_t sector_max, n_part;
    std::map<std::vector<uint32_t>, uint32_t> sectors_of;
    std::vector<std::vector<uint32_t>> partitions;
    std::vector<std::vector<double>> cycles;             // [replica][n_part] counts since the last write
    std::vector<std::map<uint32_t, uint32_t>> sector_n;  // [replica] sector -> measurements
    std::vector<uint64_t> n_measure;
    bool h5_started = false;
    Permutation(const Input& in, Engine& e, const Path& path) : Observable(in, e) {
        species = in.GetAttribute<std::string>("species");
        sector_max = in.GetAttribute<uint32_t>("sector_max", 0);
        n_part = path.n_part;
        if (n_part <= 60) {
            // ascending compositions of n (the loop of SetupPermSectors)
            const int n = (int)n_part;
            std::vector<int> a(n + 1, 0);
            int k = 1, y = n - 1;
            while (k != 0) {
                int x = a[k - 1] + 1;
                k -= 1;
                while (2 * x <= y) {
                    a[k] = x;
                    y -= x;
                    k += 1;
                }
                const int l = k + 1;
                while (x <= y) {
                    a[k] = x;
                    a[l] = y;
                    partitions.emplace_back(a.begin(), a.begin() + k + 2);
                    x += 1;
                    y -= 1;
                }
                a[k] = x + y;
                y = x + y - 1;
                partitions.emplace_back(a.begin(), a.begin() + k + 1);
            }
            for (uint32_t j = 0; j < partitions.size(); ++j) sectors_of[partitions[j]] = j;
        }
        Reset();
    }
    void Reset() {
        cycles.assign(eng.n_replicas, std::vector<double>(n_part, 0.));
        sector_n.assign(eng.n_replicas, {});
        n_measure.assign(eng.n_replicas, 0);
    }
    void DoEvent() override {
        double t = Now();
        std::vector<uint32_t> sigma(n_part);
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            check(spimc_perm_download(eng.ctx, r, sigma.data()), name.c_str());
            std::vector<uint32_t> len;
            std::vector<char> seen(n_part, 0);
            for (uint32_t p = 0; p < n_part; ++p) {
                if (seen[p]) continue;
                uint32_t c = 0, q = p;
                while (!seen[q]) {
                    seen[q] = 1;
                    q = sigma[q];
                    ++c;
                }
                len.push_back(c);
            }
            for (uint32_t c : len) cycles[r][c - 1] += 1.;
            std::sort(len.begin(), len.end());
            auto it = sectors_of.find(len);
            if (it != sectors_of.end()) sector_n[r][it->second] += 1;
            n_measure[r] += 1;
        }
        time_spent += Now() - t;
    }
};

// PathDump (path_dump_class.h:7-69): every `skip`-th Write appends the path of
// each replica to its output file; accumulates nothing.
struct PathDump : Observable {
    uint32_t skip, n_dump = 0, n_write_calls = 0;
    PathDump(const Input& in, Engine& e) : Observable(in, e) { skip = in.GetAttribute<uint32_t>("skip", 1); }
    void DoEvent() override {}
};

struct PairCorrelation : Observable {
    double r_min, r_max;
    uint32_t n_r;
    std::string species_a, species_b;
    PairCorrelation(const Input& in, Engine& e, const Path& path) : Observable(in, e) {
        which = SPIMC_OBS_GR;
        species_a = in.GetAttribute<std::string>("species_a", path.species);
        species_b = in.GetAttribute<std::string>("species_b", path.species);
        r_min = in.GetAttribute<double>("r_min", 0.);
        r_max = in.GetAttribute<double>("r_max", path.L / 2.);
        n_r = (uint32_t)in.GetAttribute<double>("n_r", 100);
    }
};

struct Loop : Event {
    int n_steps = 1;
    std::vector<std::shared_ptr<Event>> events;
    void DoEvent() override {
        // a loop over a single move is one device stage of n_steps attempts
        if (events.size() == 1 && dynamic_cast<Move*>(events[0].get())) {
            events[0]->DoEvents((uint32_t)n_steps);
            return;
        }
        for (int i = 0; i < n_steps; ++i)
            for (auto& e : events) e->DoEvent();
    }
};

}  // namespace spimc_host
