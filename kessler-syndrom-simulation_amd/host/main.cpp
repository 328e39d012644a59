// kessler_sim for the MI355X build: the reference CLI (main.cpp:16-40,
// 82-134) with the step loop (main.cpp:346-444) on the GPU.
//
//   kessler_sim [file] [duration_h] [dt_s] [out_every_s] [norad_ids] [options]
//
// Same positionals, options and output files (output/conjunctions.csv,
// final_state.csv, snapshots, breakups.csv, population.csv, profile.csv,
// trajectories.csv + moon.csv). Propagation and screening run on the device
// and object states stay in HBM; the host reads them back only for
// snapshots, tracked objects, breakups and the final state. Fragmentation
// (--breakup, --impact) runs on the host with the restated SBM, as in the
// reference. There is no CPU propagation or screening path: without a usable
// GPU the program stops with the device's message.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "breakup.hpp"
#include "collision.hpp"
#include "device_engine.hpp"
#include "orbital.hpp"

using namespace kessler;

namespace {

std::string quoted(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        if (c == '"') out += "\"\"";
        else out += c;
    }
    return out + "\"";
}

struct Options {
    std::vector<std::string> pos;
    bool screen = false, breakup = false, conj_explicit = false;
    ScreeningConfig scfg;
    BreakupConfig bcfg;
    std::string masses_path = "config/masses.csv";
    int impact_id = -1;
    double impact_mass = 10.0, impact_vrel = 10.0, impact_at_h = 0.0;
};

bool parse(int argc, char** argv, Options& o) {
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        auto next = [&]() -> const char* { return argv[++k]; };
        const bool has = k + 1 < argc;
        if (a == "--screen") o.screen = true;
        else if (a == "--breakup") o.breakup = o.screen = true;
        else if (a == "--lc" && has) o.bcfg.lc_m = std::atof(next());
        else if (a == "--masses" && has) o.masses_path = next();
        else if (a == "--seed" && has) o.bcfg.seed = std::strtoull(next(), nullptr, 10);
        else if (a == "--impact" && has) { o.impact_id = std::atoi(next()); o.breakup = o.screen = true; }
        else if (a == "--impact-mass" && has) o.impact_mass = std::atof(next());
        else if (a == "--impact-vrel" && has) o.impact_vrel = std::atof(next());
        else if (a == "--impact-at" && has) o.impact_at_h = std::atof(next());
        else if (a == "--conj-km" && has) { o.scfg.threshold_km = std::atof(next()); o.screen = o.conj_explicit = true; }
        else if (a == "--radius-scale" && has) o.scfg.radius_scale = std::atof(next());
        else if (a == "--gpu") o.screen = true;  // always on the GPU in this build
        else if (a == "--threads" && has) (void)next();  // host threads: not on the device path
        else if (a == "--min-vrel" && has) o.scfg.min_encounter_speed_km_s = std::atof(next()) * 1e-3;
        else if (a.rfind("--", 0) == 0) {
            std::cerr << "Option inconnue : " << a << "\n";
            return false;
        } else o.pos.push_back(a);
    }
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    Options opt;
    if (!parse(argc, argv, opt)) return 1;
    const std::string path = opt.pos.size() > 0 ? opt.pos[0] : "data/satellites_20260801_1000Z.txt";
    const double duration = (opt.pos.size() > 1 ? std::atof(opt.pos[1].c_str()) : 24.0) * 3600.0;
    const double dt = opt.pos.size() > 2 ? std::atof(opt.pos[2].c_str()) : 10.0;
    const double out_every = opt.pos.size() > 3 ? std::atof(opt.pos[3].c_str()) : 0.0;
    std::vector<int> track_ids;
    if (opt.pos.size() > 4) {
        std::stringstream ss(opt.pos[4]);
        std::string tok;
        while (std::getline(ss, tok, ',')) track_ids.push_back(std::atoi(tok.c_str()));
    }
    std::filesystem::create_directories(OUT_DIR);

    Simulation sim;
    int dropped = 0;
    sim.objects = load_catalog(path, &dropped);
    sim.epoch_jd = read_epoch_jd(path);
    sim.dt = dt;
    if (sim.objects.empty()) {
        std::cerr << "Aucun objet charge depuis " << path << "\n";
        return 1;
    }
    if (dropped) std::cerr << "Attention : " << dropped << " lignes ignorees (format inattendu).\n";
    MassTable masses;
    if (std::filesystem::exists(opt.masses_path)) {
        std::string err;
        if (!masses.load(opt.masses_path, &err)) {
            std::cerr << "Table des masses : " << err << "\n";
            return 1;
        }
    } else if (opt.breakup) {
        std::cout << opt.masses_path << " absent : masses par defaut.\n";
    }
    assign_masses(sim.objects, masses);
    const size_t catalog_count = sim.objects.size();
    int impact_index = -1;
    if (opt.impact_id >= 0) {
        for (size_t i = 0; i < sim.objects.size(); ++i)
            if (sim.objects[i].id == opt.impact_id) impact_index = static_cast<int>(i);
        if (impact_index < 0) {
            std::cerr << "--impact : aucun objet de NORAD " << opt.impact_id << " dans le catalogue.\n";
            return 1;
        }
    }
    if (opt.breakup && !opt.conj_explicit) {
        // Collisions only: the threshold is the largest possible radius sum
        // (main.cpp:251-260).
        double max_size_m = 0;
        for (const Object& o : sim.objects) max_size_m = std::max(max_size_m, o.size_m);
        opt.scfg.threshold_km = max_size_m * 1e-3 * opt.scfg.radius_scale;
    }

    Screener screener(opt.scfg);
    std::string message;
    if (!screener.set_gpu(true, &message)) {
        std::cerr << "GPU indisponible (" << message << ") : ce build n'a pas de chemin CPU.\n";
        return 1;
    }
    screener.attach(sim);
    sim.engine->cfg = opt.scfg;
    std::cout << sim.objects.size() << " objets charges. Duree " << duration / 3600.0 << " h, pas " << dt
              << " s\n"
              << "Epoch : " << jd_to_utc_string(sim.epoch_jd) << "\n"
              << "Lune a l'epoch : distance " << norm(sim.moon()) << " km\n";
    if (opt.screen)
        std::cout << "Detection sur GPU : " << message << "\n"
                  << "Detection : seuil " << opt.scfg.threshold_km * 1000.0 << " m"
                  << (opt.breakup && !opt.conj_explicit ? " (enveloppe de collision)" : "")
                  << ", rayons x" << opt.scfg.radius_scale << "\n";

    std::vector<State> initial;
    for (const Object& o : sim.objects) initial.push_back(o.s);
    const long steps = static_cast<long>(std::ceil(duration / dt));
    long next_out = out_every > 0 ? static_cast<long>(out_every / dt) : 0;
    long snap = 0;

    std::vector<size_t> tracked;
    std::ofstream traj, moon_out;
    if (!track_ids.empty()) {
        for (size_t i = 0; i < sim.objects.size(); ++i)
            if (std::find(track_ids.begin(), track_ids.end(), sim.objects[i].id) != track_ids.end())
                tracked.push_back(i);
        traj.open(OUT_DIR + "/trajectories.csv");
        traj << "t_s,id,name,x_km,y_km,z_km\n";
        moon_out.open(OUT_DIR + "/moon.csv");
        moon_out << "t_s,x_km,y_km,z_km\n";
        if (next_out == 0) next_out = std::max(1L, static_cast<long>(60.0 / dt));
    }
    auto record = [&](double t) {
        sim.sync_objects();
        for (size_t i : tracked) {
            const Object& o = sim.objects[i];
            traj << t << ',' << o.id << ',' << o.name << ',' << o.s.r.x << ',' << o.s.r.y << ','
                 << o.s.r.z << '\n';
        }
        const Vec3 m = moon_position(sim.epoch_jd + t / 86400.0);
        moon_out << t << ',' << m.x << ',' << m.y << ',' << m.z << '\n';
    };
    if (!tracked.empty()) record(0.0);

    std::ofstream conj_out, prof_out, breakup_out, pop_out;
    if (opt.screen) {
        conj_out.open(OUT_DIR + "/conjunctions.csv");
        conj_out << "t_s,id1,id2,name1,name2,miss_km,v_rel_km_s,collision\n";
        conj_out << std::setprecision(10);
        prof_out.open(OUT_DIR + "/profile.csv");
        prof_out << "t_h,objects,propagation_s,grid_s,pairs_s,pairs_threads_busy,breakup_s,other_s,"
                    "max_cell_objects,transfer_s\n";
    }
    Fragmentation frag(opt.bcfg);
    long long n_breakups = 0, n_catastrophic = 0, n_escaped = 0;
    if (opt.breakup) {
        breakup_out.open(OUT_DIR + "/breakups.csv");
        breakup_out << "t_s,event,target_id,target_name,projectile_id,projectile_name,catastrophic,"
                       "v_rel_km_s,energy_j_per_g,target_mass_kg,projectile_mass_kg,sbm_mass_kg,"
                       "expected_fragments,fragments,fragments_mass_kg,escaped\n";
        breakup_out << std::setprecision(8);
        pop_out.open(OUT_DIR + "/population.csv");
        pop_out << "t_s,t_days,alive,alive_catalog,alive_fragments,breakups,catastrophic,fragments_created\n";
        std::cout << "Fragmentation (SBM) : L_c = " << opt.bcfg.lc_m << " m, graine " << opt.bcfg.seed << "\n";
    }
    auto log_breakup = [&](const BreakupEvent& ev) {
        ++n_breakups;
        if (ev.catastrophic) ++n_catastrophic;
        n_escaped += ev.escaped;
        auto id_of = [&](int idx) { return idx >= 0 ? std::to_string(sim.objects[idx].id) : std::string(); };
        breakup_out << ev.t << ',' << ev.id << ',' << id_of(ev.target) << ',' << quoted(ev.target_name) << ','
                    << id_of(ev.projectile) << ',' << quoted(ev.projectile_name) << ',' << ev.catastrophic
                    << ',' << ev.v_rel_km_s << ',' << ev.energy_j_per_g << ',' << ev.target_mass_kg << ','
                    << ev.projectile_mass_kg << ',' << ev.sbm_mass_kg << ',' << ev.expected_fragments << ','
                    << ev.fragments << ',' << ev.fragments_mass_kg << ',' << ev.escaped << '\n';
        breakup_out.flush();
    };
    auto log_population = [&]() {
        sim.sync_objects();
        long alive = 0, alive_frag = 0;
        for (const Object& o : sim.objects)
            if (o.alive) {
                ++alive;
                if (o.is_fragment()) ++alive_frag;
            }
        pop_out << sim.t << ',' << sim.t / 86400.0 << ',' << alive << ',' << alive - alive_frag << ','
                << alive_frag << ',' << n_breakups << ',' << n_catastrophic << ','
                << frag.fragments_created() << std::endl;
    };
    long last_logged_hour = -1;
    if (opt.breakup) {
        log_population();
        last_logged_hour = 0;
    }
    bool impact_done = impact_index < 0;

    long long n_conj = 0, n_coll = 0, total_candidates = 0, total_refined = 0, max_co = 0;
    long long below[4] = {0, 0, 0, 0};
    const double bins[4] = {0.1, 1.0, 2.0, 5.0};
    double closest = 1e30, t_closest = 0, cell_km = 0;
    int closest_i = -1, closest_j = -1, max_cell = 0;
    double hour_loop = 0, hour_apply = 0;
    int hour_max_cell = 0;
    long profiled_hour = 0;
    std::vector<State> before;  // start-of-step states of colliding objects only
    using clock = std::chrono::steady_clock;
    const auto loop_start = clock::now();
    for (long n = 0; n < steps; ++n) {
        const auto step_start = clock::now();
        if (!impact_done && sim.t >= opt.impact_at_h * 3600.0) {
            impact_done = true;
            sim.sync_objects();
            log_breakup(frag.impact(sim, impact_index, opt.impact_mass, opt.impact_vrel));
            sim.device_dirty = true;
        }
        const double t0 = sim.t;
        if (opt.screen) screener.begin(before, sim.objects, sim.dt);
        sim.step();
        if (opt.screen) {
            ScreeningStats st;
            const std::vector<Conjunction> found = screener.finish(before, sim.objects, t0, sim.dt, &st);
            total_candidates += st.candidate_pairs;
            total_refined += st.refined_pairs;
            max_co = std::max(max_co, st.co_orbiting);
            cell_km = st.cell_km;
            max_cell = std::max(max_cell, st.max_cell_objects);
            hour_max_cell = std::max(hour_max_cell, st.max_cell_objects);
            bool any_collision = false;
            for (const Conjunction& c : found) {
                const Object& a = sim.objects[c.i];
                const Object& b = sim.objects[c.j];
                conj_out << c.t << ',' << a.id << ',' << b.id << ',' << quoted(a.name) << ',' << quoted(b.name)
                         << ',' << c.miss_km << ',' << c.v_rel_km_s << ',' << c.collision << '\n';
                ++n_conj;
                if (c.collision) {
                    ++n_coll;
                    any_collision = true;
                }
                for (int k = 0; k < 4; ++k)
                    if (c.miss_km < bins[k]) ++below[k];
                if (c.miss_km < closest) {
                    closest = c.miss_km;
                    closest_i = c.i;
                    closest_j = c.j;
                    t_closest = c.t;
                }
            }
            if (opt.breakup && any_collision) {
                const auto a0 = clock::now();
                sim.sync_objects();
                before.assign(sim.objects.size(), State{});
                std::vector<int32_t> idx;
                for (const Conjunction& c : found)
                    if (c.collision) {
                        idx.push_back(c.i);
                        idx.push_back(c.j);
                    }
                std::vector<double> prev(6 * idx.size());
                sim.engine->check(khip_get_previous(sim.engine->ctx(), static_cast<int64_t>(idx.size()),
                                                    idx.data(), prev.data()),
                                  "khip_get_previous");
                for (size_t k = 0; k < idx.size(); ++k) {
                    const double* d = prev.data() + 6 * k;
                    before[idx[k]] = State{{d[0], d[1], d[2]}, {d[3], d[4], d[5]}};
                }
                std::vector<ConjunctionRecord> recs;
                for (const Conjunction& c : found)
                    recs.push_back(ConjunctionRecord{c.i, c.j, c.t, c.miss_km, c.v_rel_km_s, c.collision});
                const size_t before_n = sim.objects.size();
                const std::vector<BreakupEvent> evs = frag.apply(sim, before, t0, sim.dt, recs);
                for (const BreakupEvent& ev : evs) log_breakup(ev);
                if (!evs.empty() || sim.objects.size() != before_n) sim.device_dirty = true;
                hour_apply += std::chrono::duration<double>(clock::now() - a0).count();
            }
        }
        hour_loop += std::chrono::duration<double>(clock::now() - step_start).count();
        if (opt.screen && static_cast<long>(sim.t / 3600.0) > profiled_hour) {
            profiled_hour = static_cast<long>(sim.t / 3600.0);
            // Propagation and screening are one device call: reported as
            // propagation_s; the host-side breakup time separately.
            prof_out << profiled_hour << ',' << sim.objects.size() << ',' << hour_loop - hour_apply << ",0,0,0,"
                     << hour_apply << ",0," << hour_max_cell << ",0" << std::endl;
            hour_loop = hour_apply = 0;
            hour_max_cell = 0;
        }
        if (opt.breakup && static_cast<long>(sim.t / 3600.0) > last_logged_hour) {
            last_logged_hour = static_cast<long>(sim.t / 3600.0);
            log_population();
        }
        if (next_out > 0 && (n + 1) % next_out == 0) {
            if (!tracked.empty()) record(sim.t);
            else {
                sim.sync_objects();
                write_snapshot(OUT_DIR + "/snapshot_" + std::to_string(snap++) + ".csv", sim.objects, sim.t);
            }
        }
    }
    const double t_loop = std::chrono::duration<double>(clock::now() - loop_start).count();
    sim.sync_objects();
    write_snapshot(OUT_DIR + "/final_state.csv", sim.objects, sim.t);
    write_snapshot_exact(OUT_DIR + "/final_state_exact.csv", sim.objects, sim.t);

    long lost = 0, destroyed = 0;
    double max_da = 0;
    int max_id = 0;
    for (size_t i = 0; i < catalog_count; ++i) {
        const Object& o = sim.objects[i];
        if (!o.alive) {
            if (norm(o.s.r) < R_EARTH) ++lost;
            else ++destroyed;
            continue;
        }
        const double da = std::fabs(elements(o.s).a - elements(initial[i]).a);
        if (da > max_da) {
            max_da = da;
            max_id = o.id;
        }
    }
    if (opt.breakup) std::cout << "Objets du catalogue detruits par collision : " << destroyed << "\n";
    std::cout << "Objets sous la surface terrestre : " << lost << "\n"
              << "Plus forte variation de demi-grand axe : " << max_da << " km (NORAD " << max_id << ")\n"
              << "Etat final ecrit dans " << OUT_DIR << "/final_state.csv\n";
    if (opt.screen) {
        std::cout << "\n--- Rapprochements (seuil " << opt.scfg.threshold_km << " km) ---\n"
                  << "Retenus : " << n_conj << "   dont collisions : " << n_coll << "\n";
        for (int k = 0; k < 4; ++k)
            if (bins[k] <= opt.scfg.threshold_km) std::cout << "  a moins de " << bins[k] << " km : " << below[k] << "\n";
        if (closest_i >= 0)
            std::cout << "Plus proche : " << closest * 1000.0 << " m entre " << sim.objects[closest_i].name << " ("
                      << sim.objects[closest_i].id << ") et " << sim.objects[closest_j].name << " ("
                      << sim.objects[closest_j].id << "), a t = " << t_closest / 3600.0 << " h\n";
        std::cout << "Paires volant de concert (< " << opt.scfg.min_encounter_speed_km_s * 1e3
                  << " m/s) sous le seuil, non traitees comme rencontres : jusqu'a " << max_co << " par pas\n"
                  << "Paires candidates : " << total_candidates / std::max(1L, steps)
                  << " par pas en moyenne (cellule " << cell_km << " km), " << total_refined
                  << " affinees au total\n"
                  << "Temps de la boucle sur GPU : " << t_loop << " s ("
                  << static_cast<double>(steps) * sim.objects.size() / t_loop << " objets-pas/s), cellule la plus peuplee : "
                  << max_cell << " objets\n";
    }
    if (opt.breakup) {
        long alive_frag = 0;
        for (size_t i = catalog_count; i < sim.objects.size(); ++i)
            if (sim.objects[i].alive) ++alive_frag;
        std::cout << "\n--- Fragmentation ---\n"
                  << "Evenements : " << n_breakups << " dont " << n_catastrophic << " catastrophiques\n"
                  << "Fragments crees : " << frag.fragments_created() << ", encore en orbite : " << alive_frag << "\n"
                  << "Fragments echappes (trajectoire hyperbolique, ecartes) : " << n_escaped << "\n";
    }
    return 0;
}
