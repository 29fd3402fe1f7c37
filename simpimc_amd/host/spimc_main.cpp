// spimc_main.cpp -- `spimc input.xml [--replicas R]`: runs a simpimc XML input
// whose hot path is covered by the engine (Framework::Run + Simulation,
// framework/framework_class.h:83-106, framework/simulation_class.h:16-45).
//
// Replicas stand in for the reference's MPI clone groups: R independent chains
// with seeds derived from <RNG seed>, advanced together on one GPU.  Block output
// goes to <output_prefix>.spimc.jsonl (one JSON object per <Write/>).
#include <chrono>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "simulation.h"

namespace spimc_host {

struct Simulation {
    Input in;
    Path path;
    std::vector<std::shared_ptr<Action>> actions;
    std::vector<std::shared_ptr<Event>> events;
    std::vector<std::shared_ptr<Observable>> observables;
    std::vector<std::shared_ptr<Move>> moves;
    Loop main_loop;
    Engine eng;
    spimc_system_desc desc{};
    std::ofstream out;
    uint32_t block = 0;
    double t_start = 0.;

    Simulation(const std::string& file, uint32_t n_replicas)
        : in(Input::Load(file)), path(in) {
        eng.n_replicas = n_replicas;
        std::memset(&desc, 0, sizeof desc);
        std::vector<std::string> action_names(2, "");
        bool have_kin = false;
        for (auto& a : in.GetChild("Actions").GetChildList("Action")) {
            const std::string type = a.GetAttribute<std::string>("type");
            std::shared_ptr<Action> act;
            if (type == "Kinetic") {
                act = std::make_shared<Kinetic>(a);
                action_names[0] = act->name;
                have_kin = true;
            } else if (type == "BarePairAction") {
                act = std::make_shared<BarePairAction>(a, path);
                action_names[1] = act->name;
            } else if (type == "DavidPairAction") {
                act = std::make_shared<DavidPairAction>(a, path);
                action_names[1] = act->name;
            } else {
                throw std::runtime_error("ERROR: Unrecognized Action, " + type);
            }
            actions.push_back(act);
        }
        if (!have_kin) throw std::runtime_error("a Kinetic action is required");
        if (action_names[1].empty()) action_names[1] = "pair";
        for (auto& m : in.GetChild("Moves").GetChildList("Move")) {
            const std::string type = m.GetAttribute<std::string>("type");
            std::shared_ptr<Move> mv;
            if (type == "Bisect")
                mv = std::make_shared<Bisect>(m, eng, desc, path);
            else if (type == "DisplaceParticle")
                mv = std::make_shared<DisplaceParticle>(m, eng, desc, path);
            else
                throw std::runtime_error("ERROR: Unrecognized Move, " + type);
            moves.push_back(mv);
            events.push_back(mv);
        }
        std::shared_ptr<StructureFactor> sk;
        std::shared_ptr<PairCorrelation> gr;
        std::shared_ptr<Energy> en;
        for (auto& o : in.GetChild("Observables").GetChildList("Observable")) {
            const std::string type = o.GetAttribute<std::string>("type");
            std::shared_ptr<Observable> ob;
            if (type == "Energy") {
                if (en) throw std::runtime_error("one Energy observable per input");
                ob = en = std::make_shared<Energy>(o, eng, action_names);
            } else if (type == "StructureFactor") {
                if (sk) throw std::runtime_error("one StructureFactor observable per input");
                ob = sk = std::make_shared<StructureFactor>(o, eng, path);
            } else if (type == "PairCorrelation") {
                if (gr) throw std::runtime_error("one PairCorrelation observable per input");
                ob = gr = std::make_shared<PairCorrelation>(o, eng, path);
            } else {
                throw std::runtime_error("ERROR: Unrecognized observable, " + type);
            }
            observables.push_back(ob);
            events.push_back(ob);
        }
        // descriptor: System, Species, k-set, then each action's tables
        desc.n_d = path.n_d;
        desc.n_bead = path.n_bead;
        desc.beta = path.beta;
        desc.L = path.L;
        desc.n_part = path.n_part;
        desc.lambda = path.lambda;
        desc.n_k = path.n_k();
        desc.k_n = path.k_n.data();
        if (desc.displace_step == 0.) desc.displace_step = path.L / 10.;
        for (auto& a : actions) a->Fill(desc, path);
        if (!desc.bisect_n_level) desc.bisect_n_level = 1;
        const uint64_t seed = in.HasChild("RNG") ? (uint64_t)in.GetChild("RNG").GetAttribute<double>("seed", 0.)
                                                 : (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
        std::vector<spimc_system_desc> descs(n_replicas, desc);
        check(spimc_ctx_create(descs.data(), n_replicas, seed, &eng.ctx), "context");
        for (uint32_t r = 0; r < n_replicas; ++r) {
            auto pos = path.InitPath(seed * 1000003ull + r);
            check(spimc_path_upload(eng.ctx, r, pos.data()), "path upload");
        }
        check(spimc_obs_configure(eng.ctx, en ? en->estimator : SPIMC_ENERGY_THERMAL, 0, sk ? sk->k_cut : 1e300,
                                  gr ? gr->r_min : 0., gr ? gr->r_max : 0., gr ? gr->n_r : 100),
              "observables");
        const std::string prefix = in.GetChild("IO").GetAttribute<std::string>("output_prefix");
        out.open(prefix + ".spimc.jsonl");
        out.precision(17);
        ReadLoop(in.GetChild("Algorithm"), main_loop);
    }

    std::shared_ptr<Event> Find(const std::string& n) {
        for (auto& e : events)
            if (e->name == n) return e;
        return nullptr;
    }

    // Loop::ReadLoop (events/loop_class.h:34-52)
    void ReadLoop(const Input& node, Loop& loop) {
        for (auto& c : node.GetChildList()) {
            const std::string t = c.GetName();
            if (t == "Move" || t == "Observable") {
                auto e = Find(c.GetAttribute<std::string>("name"));
                if (e)
                    loop.events.push_back(e);
                else
                    std::cerr << "WARNING: Event not found, " << c.GetAttribute<std::string>("name") << std::endl;
            } else if (t == "Write") {
                loop.events.push_back(std::make_shared<WriteEvent>(this));
            } else if (t == "Loop") {
                auto l = std::make_shared<Loop>();
                l->n_steps = c.GetAttribute<int>("n_step");
                ReadLoop(c, *l);
                loop.events.push_back(l);
            } else {
                std::cerr << "WARNING: Unrecognized event, " << t << std::endl;
            }
        }
    }

    // Writes::DoEvent (events/write_class.h:31-54)
    struct WriteEvent : Event {
        Simulation* sim;
        explicit WriteEvent(Simulation* s) : sim(s) { name = "write"; }
        void DoEvent() override { sim->WriteBlock(); }
    };

    void WriteBlock() {
        const double now = Now();
        out << "{\"block\": " << block << ", \"Moves\": {";
        for (size_t i = 0; i < moves.size(); ++i) {
            if (i) out << ", ";
            moves[i]->Write(out, path.n_bead);
        }
        out << "}, \"Observables\": {";
        bool first = true;
        for (auto& o : observables) {
            if (!first) out << ", ";
            first = false;
            if (auto e = dynamic_cast<Energy*>(o.get()))
                e->Write(out, path.n_bead);
            else if (dynamic_cast<StructureFactor*>(o.get()))
                WriteSk(*dynamic_cast<StructureFactor*>(o.get()));
            else if (dynamic_cast<PairCorrelation*>(o.get()))
                WriteGr(*dynamic_cast<PairCorrelation*>(o.get()));
        }
        out << "}}\n";
        out.flush();
        std::cout << "Clone #0, Block #" << block << ", Block Time: " << now - t_block << ", Total Time: " << now - t_start
                  << std::endl;
        t_block = now;
        ++block;
    }
    double t_block = 0.;
    double Now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    // StructureFactor::Write (structure_factor_class.h:69-85)
    void WriteSk(const StructureFactor& s) {
        out << "\"" << s.name << "\": {\"x\": [";
        for (uint32_t k = 0; k < path.n_k(); ++k) out << (k ? ", " : "") << path.kmag(k);
        out << "], \"y\": [";
        std::vector<double> v(std::max(path.n_k(), 1u));
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            uint64_t nm = 0;
            check(spimc_obs_read(eng.ctx, r, SPIMC_OBS_SK, v.data(), &nm), "sk");
            const double norm = (double)(nm ? nm : 1) * path.n_bead * path.n_part * path.n_part;
            out << (r ? ", [" : "[");
            for (uint32_t k = 0; k < path.n_k(); ++k) out << (k ? ", " : "") << v[k] / norm;
            out << "]";
        }
        out << "]}";
        check(spimc_obs_reset(eng.ctx, SPIMC_OBS_SK), "reset");
    }

    // PairCorrelation::Write (pair_correlation_class.h:177-210)
    void WriteGr(const PairCorrelation& g) {
        const double dr = (g.r_max - g.r_min) / (g.n_r - 1.);
        std::vector<double> x(g.n_r);
        for (uint32_t i = 0; i < g.n_r; ++i) x[i] = g.r_min + i * dr;
        out << "\"" << g.name << "\": {\"x\": [";
        for (uint32_t i = 0; i + 1 < g.n_r; ++i) {
            const double r1 = x[i], r2 = x[i + 1];
            out << (i ? ", " : "") << 0.75 * (r2 * r2 * r2 * r2 - r1 * r1 * r1 * r1) / (r2 * r2 * r2 - r1 * r1 * r1);
        }
        out << "], \"y\": [";
        std::vector<double> v(g.n_r);
        const double vol = path.L * path.L * path.L;
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            uint64_t nm = 0;
            check(spimc_obs_read(eng.ctx, r, SPIMC_OBS_GR, v.data(), &nm), "gr");
            const double norm = 0.5 * (double)(nm ? nm : 1) * path.n_part * (path.n_part - 1.) * path.n_bead / vol;
            out << (r ? ", [" : "[");
            for (uint32_t i = 0; i < g.n_r; ++i) {
                const double r1 = x[i], r2 = (i + 1 < g.n_r) ? x[i + 1] : 2. * x[i] - x[i - 1];
                const double bin_vol = 4. * M_PI / 3. * (r2 * r2 * r2 - r1 * r1 * r1);
                out << (i ? ", " : "") << v[i] / (bin_vol * norm);
            }
            out << "]";
        }
        out << "]}";
        check(spimc_obs_reset(eng.ctx, SPIMC_OBS_GR), "reset");
    }

    void Run() {
        std::cout << "Running main loop..." << std::endl;
        t_start = t_block = Now();
        main_loop.DoEvent();
    }
};

}  // namespace spimc_host

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: spimc input.xml [--replicas R]" << std::endl;
        return 2;
    }
    uint32_t R = 1;
    for (int i = 2; i + 1 < argc; ++i)
        if (std::string(argv[i]) == "--replicas") R = (uint32_t)std::atoi(argv[i + 1]);
    try {
        std::cout << "Running spimc on " << argv[1] << " with " << R << " replica(s)" << std::endl;
        spimc_host::Simulation sim(argv[1], R);
        sim.Run();
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
