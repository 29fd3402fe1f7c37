// spimc_main.cpp -- `spimc input.xml [--replicas R]`: runs a simpimc XML input
// whose hot path is covered by the engine (Framework::Run + Simulation,
// framework/framework_class.h:83-106, framework/simulation_class.h:16-45).
//
// Replicas stand in for the reference's MPI clone groups: R independent chains
// with seeds derived from <RNG seed>, advanced together on one GPU.  Block output
// goes to <output_prefix>.spimc.jsonl (one JSON object per <Write/>, all
// replicas) and, when HDF5 is available, to <output_prefix>.<replica>.h5 in the
// reference's layout (one file per clone group).
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

    Simulation(const std::string& file, uint32_t n_replicas, bool write_h5)
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
            } else if (type == "IlkkaPairAction") {
                act = std::make_shared<IlkkaPairAction>(a, path);
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
                mv = std::make_shared<Bisect>(m, eng, path);
            else if (type == "PermBisectIterative")
                mv = std::make_shared<PermBisectIterative>(m, eng, path);
            else if (type == "PermBisectTable")
                mv = std::make_shared<PermBisectTable>(m, eng, path);
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
            } else if (type == "PathDump") {
                ob = std::make_shared<PathDump>(o, eng);
            } else if (type == "Permutation") {
                ob = std::make_shared<Permutation>(o, eng, path);
            } else if (type == "Time") {
                ob = std::make_shared<TimeObs>(o, eng);
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
        // one set of bisection splines per replica: the Bisect move's level (else
        // the permuting one's) seeds the descriptor; all must share n_images
        uint32_t plain_level = 0, perm_level = 0;
        bool images_set = false;
        for (auto& m : moves)
            if (auto* b = dynamic_cast<Bisect*>(m.get())) {
                uint32_t& lv = dynamic_cast<PermBisectIterative*>(m.get()) ? perm_level : plain_level;
                if (!lv) lv = b->n_level0;
                if (images_set && desc.bisect_n_images != b->n_images)
                    throw std::runtime_error("ERROR: bisection moves with different n_images are outside this engine");
                desc.bisect_n_images = b->n_images;
                images_set = true;
            }
        desc.bisect_n_level = plain_level ? plain_level : (perm_level ? perm_level : 1);
        const uint64_t seed = in.HasChild("RNG") ? (uint64_t)in.GetChild("RNG").GetAttribute<double>("seed", 0.)
                                                 : (uint64_t)std::chrono::system_clock::now().time_since_epoch().count();
        std::vector<spimc_system_desc> descs(n_replicas, desc);
        check(spimc_ctx_create(descs.data(), n_replicas, seed, &eng.ctx), "context");
        for (uint32_t r = 0; r < n_replicas; ++r) {
            std::vector<uint32_t> sigma;
            auto pos = path.InitPath(seed * 1000003ull + r, r, sigma);
            check(spimc_path_upload(eng.ctx, r, pos.data()), "path upload");
            check(spimc_perm_upload(eng.ctx, r, sigma.data()), "permutation upload");
        }
        for (auto& m : moves)
            if (auto* pb = dynamic_cast<PermBisectIterative*>(m.get())) pb->Configure();
        check(spimc_obs_configure(eng.ctx, en ? en->estimator : SPIMC_ENERGY_THERMAL, 0, sk ? sk->k_cut : 1e300,
                                  gr ? gr->r_min : 0., gr ? gr->r_max : 0., gr ? gr->n_r : 100),
              "observables");
        const std::string prefix = in.GetChild("IO").GetAttribute<std::string>("output_prefix");
        out.open(prefix + ".spimc.jsonl");
        out.precision(17);
        if (write_h5) {
#if SPIMC_HAVE_HDF5
            OpenH5(prefix, file, seed);
            energy_started_.assign(eng.n_replicas, false);
#else
            std::cerr << "WARNING: built without HDF5; block output goes to " << prefix << ".spimc.jsonl only" << std::endl;
#endif
        }
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
            WriteMove(*moves[i]);
        }
        out << "}, \"Observables\": {";
        bool first = true;
        for (auto& o : observables) {
            if (!first) out << ", ";
            first = false;
            if (auto e = dynamic_cast<Energy*>(o.get()))
                WriteEnergy(*e);
            else if (auto s = dynamic_cast<StructureFactor*>(o.get()))
                WriteSk(*s);
            else if (auto g = dynamic_cast<PairCorrelation*>(o.get()))
                WriteGr(*g);
            else if (auto d = dynamic_cast<PathDump*>(o.get()))
                WritePathDump(*d);
            else if (auto t = dynamic_cast<TimeObs*>(o.get()))
                WriteTime(*t);
            else if (auto pm = dynamic_cast<Permutation*>(o.get()))
                WritePermutation(*pm);
        }
        out << "}}\n";
        out.flush();
#if SPIMC_HAVE_HDF5
        for (auto& f : h5) f->Flush();
#endif
        std::cout << "Clone #0, Block #" << block << ", Block Time: " << now - t_block << ", Total Time: " << now - t_start
                  << std::endl;
        t_block = now;
        ++block;
    }
    double t_block = 0.;
    double Now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
    }

    // Move::Write (move_class.h:80-94); Bisect::Write first records the block's
    // level and adapts it (bisect_class.h:193-211)
    void WriteMove(Move& m) {
        const auto b = m.Block();
        auto* bis = dynamic_cast<Bisect*>(&m);
        auto* pb = dynamic_cast<PermBisectIterative*>(&m);
        // PermBisect::Write (perm_bisect_class.h:119-132): the block's counts per
        // cycle length, before the bisection level and Move::Write
        if (pb) pb->PermBlock();
        std::vector<uint32_t> level;
        if (bis) {
            bis->n_level.resize(eng.n_replicas, bis->n_level0);
            level = bis->n_level;
            bis->Adapt(b);
        }
        out << "\"" << m.name << "\": [";
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            char s[200];
            snprintf(s, sizeof s, "%s{\"n_attempt\": %llu, \"n_accept\": %llu, \"x\": %.10g", r ? ", " : "",
                     (unsigned long long)b[r].n_attempt, (unsigned long long)b[r].n_accept,
                     b[r].n_attempt ? (double)b[r].n_accept / b[r].n_attempt : 0.);
            out << s;
            if (pb) {
                for (int w = 0; w < 2; ++w) {
                    out << (w ? ", \"perm_accept\": [" : ", \"perm_attempt\": [");
                    const auto& v = w ? pb->perm_acc[r] : pb->perm_att[r];
                    for (size_t c = 0; c < v.size(); ++c) out << (c ? ", " : "") << v[c];
                    out << "]";
                }
            }
            if (bis && bis->adaptive) out << ", \"n_level\": " << level[r];
            out << "}";
        }
        out << "]";
#if SPIMC_HAVE_HDF5
        if (pb)
            for (uint32_t r = 0; r < h5.size(); ++r) {
                const std::string g = "Moves/" + m.name + "/";
                if (!pb->h5_started) h5[r]->Int(g + "n_perm_type", pb->NPermType());
                h5[r]->AppendInts(g + "perm_attempt", std::vector<long long>(pb->perm_att[r].begin(), pb->perm_att[r].end()));
                h5[r]->AppendInts(g + "perm_accept", std::vector<long long>(pb->perm_acc[r].begin(), pb->perm_acc[r].end()));
            }
        if (pb) pb->h5_started = true;
        if (bis && bis->adaptive)
            for (uint32_t r = 0; r < h5.size(); ++r) h5[r]->AppendInt("Moves/" + m.name + "/n_level", level[r]);
#endif
#if SPIMC_HAVE_HDF5
        for (uint32_t r = 0; r < h5.size(); ++r) {
            const std::string g = "Moves/" + m.name + "/";
            h5[r]->AppendInt(g + "n_attempt", (long long)b[r].n_attempt);
            h5[r]->AppendInt(g + "n_accept", (long long)b[r].n_accept);
            const double x = (1. * b[r].n_accept) / (1. * b[r].n_attempt);  // 0/0 is NaN, as written by the reference
            h5[r]->Append(g + "x", &x, 0);
        }
#endif
    }

    // Energy::Write (energy_class.h:232-256), without sectors and potentials
    void WriteEnergy(Energy& e) {
        const auto b = e.Block(path.n_bead);
        out << "\"" << e.name << "\": [";
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            char s[200];
            snprintf(s, sizeof s, "%s{\"total\": %.15g, \"%s\": %.15g, \"%s\": %.15g, \"n_measure\": %llu}", r ? ", " : "",
                     b[r].kinetic + b[r].pair, e.action_names[0].c_str(), b[r].kinetic, e.action_names[1].c_str(), b[r].pair,
                     (unsigned long long)b[r].n_measure);
            out << s;
        }
        out << "]";
#if SPIMC_HAVE_HDF5
        for (uint32_t r = 0; r < h5.size(); ++r) {
            if (!b[r].n_measure) continue;
            const std::string g = "Observables/" + e.name + "/";
            const double vals[3] = {b[r].kinetic + b[r].pair, b[r].kinetic, b[r].pair};
            const std::string names[3] = {"total", e.action_names[0], e.action_names[1]};
            for (int i = 0; i < 3; ++i) {
                if (!energy_started_[r]) h5[r]->String(g + names[i] + "/data_type", "scalar");
                h5[r]->Append(g + names[i] + "/x", &vals[i], 0);
            }
            energy_started_[r] = true;
        }
#endif
    }
    std::vector<bool> energy_started_;

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
            for (uint32_t k = 0; k < path.n_k(); ++k) v[k] /= norm;
            out << (r ? ", [" : "[");
            for (uint32_t k = 0; k < path.n_k(); ++k) out << (k ? ", " : "") << v[k];
            out << "]";
#if SPIMC_HAVE_HDF5
            if (nm && r < h5.size()) h5[r]->Append("Observables/" + s.name + "/y", v.data(), path.n_k());
#endif
        }
        out << "]}";
        check(spimc_obs_reset(eng.ctx, SPIMC_OBS_SK), "reset");
    }

    // PairCorrelation bin centres (pair_correlation_class.h:64-74): the
    // r^3-weighted mean radius of each of the n_r - 1 shells
    static std::vector<double> GrCentres(const PairCorrelation& g) {
        const double dr = (g.r_max - g.r_min) / (g.n_r - 1.);
        std::vector<double> rs(g.n_r - 1);
        for (uint32_t i = 0; i + 1 < g.n_r; ++i) {
            const double r1 = g.r_min + i * dr, r2 = g.r_min + (i + 1) * dr;
            rs[i] = 0.75 * (r2 * r2 * r2 * r2 - r1 * r1 * r1 * r1) / (r2 * r2 * r2 - r1 * r1 * r1);
        }
        return rs;
    }

    // PairCorrelation::Write (pair_correlation_class.h:89-124)
    void WriteGr(const PairCorrelation& g) {
        const double dr = (g.r_max - g.r_min) / (g.n_r - 1.);
        std::vector<double> x(g.n_r);
        for (uint32_t i = 0; i < g.n_r; ++i) x[i] = g.r_min + i * dr;
        out << "\"" << g.name << "\": {\"x\": [";
        const auto rs = GrCentres(g);
        for (size_t i = 0; i < rs.size(); ++i) out << (i ? ", " : "") << rs[i];
        out << "], \"y\": [";
        std::vector<double> v(g.n_r);
        const double vol = path.L * path.L * path.L;
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            uint64_t nm = 0;
            check(spimc_obs_read(eng.ctx, r, SPIMC_OBS_GR, v.data(), &nm), "gr");
            const double norm = 0.5 * (double)(nm ? nm : 1) * path.n_part * (path.n_part - 1.) * path.n_bead / vol;
            for (uint32_t i = 0; i < g.n_r; ++i) {
                const double r1 = x[i], r2 = (i + 1 < g.n_r) ? x[i + 1] : 2. * x[i] - x[i - 1];
                const double bin_vol = 4. * M_PI / 3. * (r2 * r2 * r2 - r1 * r1 * r1);
                v[i] /= bin_vol * norm;
            }
            out << (r ? ", [" : "[");
            for (uint32_t i = 0; i < g.n_r; ++i) out << (i ? ", " : "") << v[i];
            out << "]";
#if SPIMC_HAVE_HDF5
            if (nm && r < h5.size()) h5[r]->Append("Observables/" + g.name + "/y", v.data(), g.n_r);
#endif
        }
        out << "]}";
        check(spimc_obs_reset(eng.ctx, SPIMC_OBS_GR), "reset");
    }

    // Time::Write (time_class.h:38-67): block time and per-event times (seconds),
    // each an extendable scalar with data_type "scalar"; then every event's
    // time_spent is cleared (Time::Reset)
    void WriteTime(TimeObs& t) {
        const double now = Now(), total = now - t.start;
        out << "\"" << t.name << "\": {\"block\": " << total;
        for (auto& e : events) out << ", \"" << e->name << "\": " << e->time_spent;
        out << "}";
#if SPIMC_HAVE_HDF5
        for (auto& f : h5) {
            const std::string g = "Observables/" + t.name + "/";
            if (!time_started_) {
                f->String(g + "block/data_type", "scalar");
                for (auto& e : events) f->String(g + e->name + "/data_type", "scalar");
            }
            f->Append(g + "block/x", &total, 0);
            for (auto& e : events) f->Append(g + e->name + "/x", &e->time_spent, 0);
        }
#endif
        time_started_ = true;
        for (auto& e : events) e->time_spent = 0.;
        t.start = now;
    }
    bool time_started_ = false;

    // Permutation::Write (permutation_class.h:76-118): per replica the block's
    // (sector, count) pairs in sector order, appended row by row to sectors/y,
    // and the cycle-length histogram normalised by the number of cycles to
    // cycles/y; the constructor's datasets go out with the first write.
    void WritePermutation(Permutation& pm) {
        out << "\"" << pm.name << "\": [";
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            double n_cyc = 0.;
            for (double c : pm.cycles[r]) n_cyc += c;
            out << (r ? ", " : "") << "{\"n_measure\": " << pm.n_measure[r] << ", \"cycles\": [";
            for (uint32_t c = 0; c < pm.n_part; ++c)
                out << (c ? ", " : "") << (n_cyc > 0. ? pm.cycles[r][c] / n_cyc : 0.);
            out << "], \"sectors\": [";
            bool f = true;
            for (auto& kv : pm.sector_n[r]) {
                out << (f ? "" : ", ") << "[" << kv.first << ", " << kv.second << "]";
                f = false;
            }
            out << "]}";
        }
        out << "]";
#if SPIMC_HAVE_HDF5
        const std::string g = "Observables/" + pm.name + "/";
        for (uint32_t r = 0; r < h5.size(); ++r) {
            if (!pm.h5_started) {
                h5[r]->String(g + "species", pm.species);
                h5[r]->Int(g + "sector_max", pm.sector_max);
                h5[r]->String(g + "sectors/data_type", "pairs");
                std::vector<long long> x(pm.partitions.size());
                std::vector<long long> poss((size_t)pm.n_part * pm.partitions.size(), 0);
                for (size_t j = 0; j < pm.partitions.size(); ++j) {
                    x[j] = (long long)j;
                    for (uint32_t len : pm.partitions[j]) poss[(size_t)(len - 1) * pm.partitions.size() + j] += 1;
                }
                h5[r]->Ints(g + "sectors/x", x);
                h5[r]->Ints(g + "sectors/possPerms", poss, {pm.n_part, (hsize_t)pm.partitions.size()});
                h5[r]->String(g + "cycles/data_type", "histogram");
                std::vector<long long> cx(pm.n_part);
                for (uint32_t c = 0; c < pm.n_part; ++c) cx[c] = c + 1;
                h5[r]->Ints(g + "cycles/x", cx);
            }
            if (pm.n_measure[r] == 0) continue;
            for (auto& kv : pm.sector_n[r]) h5[r]->AppendInts(g + "sectors/y", {(long long)kv.first, (long long)kv.second});
            double n_cyc = 0.;
            for (double c : pm.cycles[r]) n_cyc += c;
            std::vector<double> y(pm.n_part);
            for (uint32_t c = 0; c < pm.n_part; ++c) y[c] = pm.cycles[r][c] / n_cyc;
            h5[r]->Append(g + "cycles/y", y.data(), pm.n_part);
        }
#endif
        pm.h5_started = true;
        pm.Reset();
    }

    // PathDump::Write (path_dump_class.h:29-66): positions [dump][particle][bead][d]
    // (the reference's cube(n_d, n_bead, n_part) per dump, first index fastest)
    // and permutation [dump][particle][2] (previous and next particle of the
    // end beads, the identity for Boltzmann paths)
    void WritePathDump(PathDump& d) {
        const bool dump = !(d.n_write_calls % d.skip);
        if (dump) d.n_dump += 1;
        out << "\"" << d.name << "\": {\"n_dump\": " << d.n_dump << "}";
#if SPIMC_HAVE_HDF5
        if (dump && !h5.empty()) {
            const uint32_t N = path.n_part, M = path.n_bead;
            std::vector<double> r((size_t)3 * M * N), pd((size_t)N * M * 3), perm((size_t)2 * N);
            std::vector<uint32_t> sigma(N);
            const std::string g = "Observables/" + d.name + "/" + path.species + "/";
            for (uint32_t rep = 0; rep < h5.size(); ++rep) {
                check(spimc_path_download(eng.ctx, rep, r.data()), "path download");
                check(spimc_perm_download(eng.ctx, rep, sigma.data()), "permutation download");
                for (uint32_t p = 0; p < N; ++p) {
                    perm[2 * sigma[p]] = p;  // previous particle of sigma[p]'s first bead
                    perm[2 * p + 1] = sigma[p];
                }
                for (uint32_t p = 0; p < N; ++p)
                    for (uint32_t b = 0; b < M; ++b)
                        for (uint32_t x = 0; x < 3; ++x) pd[((size_t)p * M + b) * 3 + x] = r[((size_t)x * M + b) * N + p];
                h5[rep]->SetInt(g + "n_dump", d.n_dump);
                h5[rep]->Append(g + "positions", pd.data(), {N, M, 3});
                h5[rep]->Append(g + "permutation", perm.data(), {N, 2});
            }
        }
#endif
        d.n_write_calls += 1;
    }

#if SPIMC_HAVE_HDF5
    std::vector<std::unique_ptr<H5Out>> h5;

    // One file per replica, <output_prefix>.<replica>.h5, as the reference
    // writes one per clone group (framework_class.h:68-80); the set-up
    // records of Framework, Simulation, Action, Move and Observable.
    void OpenH5(const std::string& prefix, const std::string& in_file, uint64_t seed) {
        std::ifstream f(in_file);
        std::stringstream contents;
        contents << f.rdbuf();
        for (uint32_t r = 0; r < eng.n_replicas; ++r) {
            h5.push_back(std::make_unique<H5Out>(prefix + "." + std::to_string(r) + ".h5"));
            H5Out& o = *h5.back();
            o.String("Input/file_name", in_file);
            o.String("Input/contents", contents.str());
            o.Int("RNG/seed", (long long)seed);
            o.Group("Actions");
            for (auto& a : actions) {
                const std::string g = "Actions/" + a->name + "/";
                o.String(g + "type", a->type);
                o.Int(g + "n_images", a->n_images);
                o.Int(g + "max_level", a->max_level);
                if (auto p = dynamic_cast<PairActionBase*>(a.get())) {
                    o.Int(g + "n_order", p->n_order);
                    o.String(g + "species_a", p->species_a);
                    o.String(g + "species_b", p->species_b);
                    o.Int(g + "use_long_range", p->use_long_range);
                    if (p->use_long_range) o.Double(g + "k_cut", p->k_cut);
                    o.String(g + "file", p->file);
                }
            }
            o.Group("Moves");
            for (auto& m : moves) {
                o.String("Moves/" + m->name + "/type", m->type);
                o.String("Moves/" + m->name + "/data_type", "scalar");
            }
            o.Group("Observables");
            for (auto& ob : observables) {
                const std::string g = "Observables/" + ob->name + "/";
                o.String(g + "type", ob->type);
                const bool hist = dynamic_cast<StructureFactor*>(ob.get()) || dynamic_cast<PairCorrelation*>(ob.get());
                o.String(g + "data_type", hist ? "histogram" : "none");
                if (auto s = dynamic_cast<StructureFactor*>(ob.get())) {
                    o.String(g + "species_a", s->species_a);
                    o.String(g + "species_b", s->species_b);
                    o.Double(g + "k_cut", s->k_cut);
                    std::vector<double> mags(path.n_k());
                    for (uint32_t k = 0; k < path.n_k(); ++k) mags[k] = path.kmag(k);
                    o.Doubles(g + "x", mags);
                } else if (auto gr = dynamic_cast<PairCorrelation*>(ob.get())) {
                    o.String(g + "species_a", gr->species_a);
                    o.String(g + "species_b", gr->species_b);
                    o.Double(g + "r_min", gr->r_min);
                    o.Double(g + "r_max", gr->r_max);
                    o.Int(g + "n_r", gr->n_r);
                    o.Doubles(g + "x", GrCentres(*gr));
                }
            }
            o.Flush();
        }
    }
#endif

    void Run() {
        std::cout << "Running main loop..." << std::endl;
        t_start = t_block = Now();
        main_loop.DoEvent();
    }
};

}  // namespace spimc_host

// `spimc --table-to-h5 in out [--transposed]`: writes a pair-action table
// (text container or HDF5) as an HDF5 file with the same dataset names.  The
// David u_kj / du_kj_dbeta cubes become 3-D datasets [n_tau][n_grid][n_val] and
// the Ilkka u_xy / du_xy matrices 2-D datasets [n_y][n_x] (the reference's
// memory order), or [n_val][n_grid][n_tau] and [n_x][n_y] with --transposed;
// everything else keeps its shape.
static int TableToH5(const std::string& in_path, const std::string& out_path, bool tau_fastest) {
#if SPIMC_HAVE_HDF5
    spimc_host::TableFile t;
    t.Load(in_path);
    const auto& d = t.Data();
    spimc_host::H5Out o(out_path);
    for (auto& kv : d.strings) o.String(kv.first, kv.second);
    for (auto& kv : d.data) {
        const std::string& name = kv.first;
        const std::vector<double>& v = kv.second;
        std::string pre;
        for (const char* p : {"u_kj_", "du_kj_dbeta_"})
            if (name.rfind(p, 0) == 0 && name.size() > 5 && name.compare(name.size() - 5, 5, "/data") == 0)
                pre = "u_kj_" + name.substr(std::strlen(p), name.size() - 5 - std::strlen(p));
        if (!pre.empty() && t.Has(pre + "/grid/n_grid_points") && t.Has(pre + "/taus")) {
            const size_t n_grid = (size_t)t.Scalar(pre + "/grid/n_grid_points"), n_tau = t.Read(pre + "/taus").size();
            const size_t n_val = v.size() / (n_grid * n_tau);
            const std::vector<double> c = t.Cube(name, n_val, n_grid, n_tau);  // first index fastest
            if (!tau_fastest) {
                o.Doubles(name, c, {n_tau, n_grid, n_val});
            } else {
                std::vector<double> w(v.size());
                for (size_t i = 0; i < n_val; ++i)
                    for (size_t g = 0; g < n_grid; ++g)
                        for (size_t k = 0; k < n_tau; ++k) w[(i * n_grid + g) * n_tau + k] = c[i + n_val * (g + n_grid * k)];
                o.Doubles(name, w, {n_val, n_grid, n_tau});
            }
        } else if ((name == "u/off_diag/u_xy" || name == "du/off_diag/du_xy") &&
                   t.Has(name.substr(0, name.find('/')) + "/off_diag/n_x")) {
            const std::string g = name.substr(0, name.find('/'));
            const size_t n_x = (size_t)t.Scalar(g + "/off_diag/n_x"), n_y = (size_t)t.Scalar(g + "/off_diag/n_y");
            const std::vector<double> m = t.Matrix(name, n_x, n_y);  // x fastest
            if (!tau_fastest) {
                o.Doubles(name, m, {n_y, n_x});
            } else {
                std::vector<double> w(m.size());
                for (size_t i = 0; i < n_x; ++i)
                    for (size_t j = 0; j < n_y; ++j) w[i * n_y + j] = m[i + n_x * j];
                o.Doubles(name, w, {n_x, n_y});
            }
        } else if (d.ints.count(name)) {
            std::vector<long long> iv(v.begin(), v.end());
            o.Ints(name, iv);
        } else {
            const auto& dm = d.dims.at(name);
            o.Doubles(name, v, std::vector<hsize_t>(dm.begin(), dm.end()));
        }
    }
    std::cout << "wrote " << out_path << ": " << d.data.size() + d.strings.size() << " datasets" << std::endl;
    return 0;
#else
    (void)in_path;
    (void)out_path;
    (void)tau_fastest;
    std::cerr << "ERROR: built without HDF5" << std::endl;
    return 1;
#endif
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::cerr << "usage: spimc input.xml [--replicas R] [--no-h5]\n"
                     "       spimc --table-to-h5 table_in table_out.h5 [--transposed]"
                  << std::endl;
        return 2;
    }
    if (std::string(argv[1]) == "--table-to-h5") {
        if (argc < 4) {
            std::cerr << "usage: spimc --table-to-h5 table_in table_out.h5 [--transposed]" << std::endl;
            return 2;
        }
        try {
            return TableToH5(argv[2], argv[3], argc > 4 && std::string(argv[4]) == "--transposed");
        } catch (const std::exception& e) {
            std::cerr << "ERROR: " << e.what() << std::endl;
            return 1;
        }
    }
    uint32_t R = 1;
    bool h5 = true;
    for (int i = 2; i < argc; ++i) {
        if (std::string(argv[i]) == "--replicas" && i + 1 < argc) R = (uint32_t)std::atoi(argv[i + 1]);
        if (std::string(argv[i]) == "--no-h5") h5 = false;
    }
    try {
        std::cout << "Running spimc on " << argv[1] << " with " << R << " replica(s)" << std::endl;
        spimc_host::Simulation sim(argv[1], R, h5);
        sim.Run();
    } catch (const std::exception& e) {
        std::cerr << "ERROR: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
