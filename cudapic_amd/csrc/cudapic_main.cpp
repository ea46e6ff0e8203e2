// cudapic_main.cpp -- C++ host of the engine: the reference's run interface
// (Rust-CUDA/src/main.rs:1611-1878) on the C ABI of include/cudapic_amd.h.
//
//   cudapic <num_cycles> [--measure <measurement_cycles>] [options]
//
// Writes results/<timestamp>/ like the reference (main.rs:1498-1609):
//   NNNN_<ts>_electrons.csv, NNNN_<ts>_ions.csv   x,vx,vy,vz of the live particles
//   particle_growth_<ts>.csv                       step,n_e,n_i every 100 cycles
//   density_avg_<ts>.csv, eepf_<ts>.csv            with --measure
// and, with --measure, ifed_<ts>.csv and info_<ts>.txt (electrode particle
// fluxes and mean ion energies, the quantities of C/eduPIC.cc:1002-1008);
// with --measure and --xt BIN also eduPIC's spatio-temporal maps pot_xt.dat,
// efield_xt.dat, ne_xt.dat, ni_xt.dat, je_xt.dat, ji_xt.dat, powere_xt.dat,
// poweri_xt.dat, meanee_xt.dat, meanei_xt.dat, ioniz_xt.dat (N_G rows of
// N_T / BIN columns, C/eduPIC.cc:847-919; eduPIC's BIN is 20).
//
// Options beyond the reference's compile-time constants:
//   --path fused|reference   --precision fp32|comp|fp64   --no-graph
//   --weight W  --voltage V  --pressure P  --ng N  --capacity N  --n-init N
//   --thermal-init (cudaPIC's thermal start, main.rs:1361-1376; default is
//   eduPIC's start at rest, C/eduPIC.cc:283-294)   --seed S
//   --load FILE --save FILE (binary checkpoint)    --out DIR
//   --import-edupic FILE / --export-edupic FILE: eduPIC's picdata.bin
//   (C/eduPIC.cc:726-781: Time, cycles_done, N_e as f64, then x, vx, vy, vz
//   of the electrons as f64, N_i and the ions likewise); --int-headers for
//   the C++ eduPIC's variant with int32 cycles_done, N_e and N_i
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/cudapic_amd.h"

namespace {

constexpr int kCheckpointCycles = 100;  // main.rs:110

void die(const char *what, int rc) {
    std::fprintf(stderr, ">> cudapic: %s failed (status %d): %s\n", what, rc, pic_last_error());
    std::exit(1);
}
#define CK(call)                    \
    do {                            \
        int rc_ = (call);           \
        if (rc_ != PIC_OK) die(#call, rc_); \
    } while (0)

void usage() {
    std::fprintf(stderr,
                 "usage: cudapic <num_cycles> [--measure <measurement_cycles>] [--path fused|reference]\n"
                 "       [--precision fp32|comp|fp64] [--no-graph] [--weight W] [--voltage V] [--pressure P]\n"
                 "       [--ng N] [--capacity N] [--n-init N] [--thermal-init] [--seed S]\n"
                 "       [--load FILE] [--save FILE] [--out DIR]\n"
                 "       [--import-edupic FILE] [--export-edupic FILE] [--int-headers] [--xt BIN]\n");
}

// xoshiro128+ for the seeded initial positions (the reference draws them from
// an unseeded generator, main.rs:1362)
struct Xo {
    uint32_t s[4];
    uint32_t next() {
        const uint32_t r = s[0] + s[3], t = s[1] << 9;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = (s[3] << 11) | (s[3] >> 21);
        return r;
    }
    double uniform() { return (double)(next() >> 8) * (1.0 / 16777216.0); }
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-12) u1 = 1e-12;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

std::string timestamp() {
    std::time_t t = std::time(nullptr);
    char buf[64];
    std::strftime(buf, sizeof(buf), "%Y-%m-%d_%H-%M-%S", std::localtime(&t));
    return buf;
}

void mkdirs(const std::string &path) {
    std::string cur;
    for (size_t i = 0; i < path.size(); ++i) {
        cur += path[i];
        if (path[i] == '/' || i + 1 == path.size()) mkdir(cur.c_str(), 0755);
    }
}

}  // namespace

// eduPIC's particle data file (C/eduPIC.cc:726-781)
struct EduState {
    double time = 0.0, cycles = 0.0;
    std::vector<float> sp[2][4];  // electrons, ions: x, vx, vy, vz
};

static bool read_edupic(const std::string &path, bool int_headers, EduState &st) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    auto header = [&](double &v) {
        if (int_headers) {
            int32_t i = 0;
            if (std::fread(&i, 4, 1, f) != 1) return false;
            v = i;
            return true;
        }
        return std::fread(&v, 8, 1, f) == 1;
    };
    bool ok = std::fread(&st.time, 8, 1, f) == 1 && header(st.cycles);
    for (int sp = 0; ok && sp < 2; ++sp) {
        double n = 0.0;
        ok = header(n) && n >= 0.0 && n < 4.3e9;
        std::vector<double> buf(ok ? (size_t)n : 0);
        for (int c = 0; ok && c < 4; ++c) {
            ok = std::fread(buf.data(), 8, buf.size(), f) == buf.size();
            st.sp[sp][c].assign(buf.begin(), buf.end());
        }
    }
    std::fclose(f);
    return ok;
}

static bool write_edupic(const std::string &path, bool int_headers, const EduState &st) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    auto header = [&](double v) {
        if (int_headers) {
            const int32_t i = (int32_t)v;
            return std::fwrite(&i, 4, 1, f) == 1;
        }
        return std::fwrite(&v, 8, 1, f) == 1;
    };
    bool ok = std::fwrite(&st.time, 8, 1, f) == 1 && header(st.cycles);
    for (int sp = 0; ok && sp < 2; ++sp) {
        ok = header((double)st.sp[sp][0].size());
        for (int c = 0; ok && c < 4; ++c) {
            std::vector<double> buf(st.sp[sp][c].begin(), st.sp[sp][c].end());
            ok = std::fwrite(buf.data(), 8, buf.size(), f) == buf.size();
        }
    }
    return std::fclose(f) == 0 && ok;
}

int main(int argc, char **argv) {
    std::printf(">> cudaPIC (MI355X): starting...\n");
    if (argc < 2) {
        usage();
        return 1;
    }
    const long num_cycles = std::strtol(argv[1], nullptr, 10);
    if (num_cycles < 0) {
        usage();
        return 1;
    }
    bool measure = false, graph = true, thermal = false;
    long measurement_cycles = 0;
    std::string path = "fused", precision = "fp64", load, save, out_root = "results", import_ed, export_ed;
    bool int_headers = false;
    uint32_t xt_bin = 0;
    double weight = 7.0e4, voltage = 250.0, pressure = 10.0;
    int n_g = 400;
    long capacity = 120000, n_init = 1000;
    uint32_t seed = 0;
    for (int i = 2; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) {
                usage();
                std::exit(1);
            }
            return argv[++i];
        };
        if (a == "--measure") {
            measure = true;
            if (i + 1 < argc && argv[i + 1][0] != '-') measurement_cycles = std::strtol(argv[++i], nullptr, 10);
        } else if (a == "--path") path = val();
        else if (a == "--precision") precision = val();
        else if (a == "--no-graph") graph = false;
        else if (a == "--weight") weight = std::atof(val().c_str());
        else if (a == "--voltage") voltage = std::atof(val().c_str());
        else if (a == "--pressure") pressure = std::atof(val().c_str());
        else if (a == "--ng") n_g = std::atoi(val().c_str());
        else if (a == "--capacity") capacity = std::strtol(val().c_str(), nullptr, 10);
        else if (a == "--n-init") n_init = std::strtol(val().c_str(), nullptr, 10);
        else if (a == "--thermal-init") thermal = true;
        else if (a == "--seed") seed = (uint32_t)std::strtoul(val().c_str(), nullptr, 10);
        else if (a == "--load") load = val();
        else if (a == "--save") save = val();
        else if (a == "--import-edupic") import_ed = val();
        else if (a == "--export-edupic") export_ed = val();
        else if (a == "--int-headers") int_headers = true;
        else if (a == "--xt") xt_bin = (uint32_t)std::strtoul(val().c_str(), nullptr, 10);
        else if (a == "--out") out_root = val();
        else {
            usage();
            return 1;
        }
    }
    if (measurement_cycles > num_cycles) measurement_cycles = num_cycles;
    const long measurement_start = num_cycles - measurement_cycles;
    std::printf(">> cudaPIC: num_cycles = %ld, measure = %s", num_cycles, measure ? "true" : "false");
    if (measure) std::printf(", measurement_start_cycle = %ld", measurement_start);
    std::printf("\n");

    auto t_init = std::chrono::steady_clock::now();
    pic_sim_config_t cfg;
    pic_sim_config_default(&cfg);
    pic_params_default(&cfg.params);
    cfg.params.n_g = n_g;
    cfg.params.weight = weight;
    cfg.params.v0 = voltage;
    cfg.params.pressure = pressure;
    cfg.params.precision = precision == "fp32" ? PIC_PREC_FP32 : precision == "comp" ? PIC_PREC_COMP : PIC_PREC_FP64;
    CK(pic_params_derive(&cfg.params));
    cfg.capacity = (uint32_t)capacity;
    cfg.path = path == "reference" ? PIC_PATH_REFERENCE : PIC_PATH_FUSED;
    cfg.use_graph = graph ? 1 : 0;
    cfg.seed_e[0] += seed;
    cfg.seed_i[0] += seed;
    pic_sim_t *sim = nullptr;
    CK(pic_sim_create(&cfg, &sim));
    const pic_params_t &p = cfg.params;
    double ed_time = 0.0, ed_cycles = 0.0;
    if (!load.empty()) {
        CK(pic_sim_load(sim, load.c_str()));
        std::printf(">> cudaPIC: state loaded from %s\n", load.c_str());
    } else if (!import_ed.empty()) {
        EduState st;
        if (!read_edupic(import_ed, int_headers, st)) {
            std::fprintf(stderr, ">> cudaPIC: cannot read eduPIC particle data %s\n", import_ed.c_str());
            return 1;
        }
        ed_time = st.time;
        ed_cycles = st.cycles;
        for (int sp = 0; sp < 2; ++sp) {
            const auto &v = st.sp[sp];
            const uint32_t n = (uint32_t)v[0].size();
            CK(pic_sim_upload(sim, sp, v[0].data(), v[1].data(), v[2].data(), v[3].data(), n));
        }
        // the step of the RF cycle the saved time falls on
        const long step = std::lround(st.time / p.dt_e);
        CK(pic_sim_set_time(sim, (uint32_t)(((step % p.n_t) + p.n_t) % p.n_t)));
        std::printf(">> cudaPIC: eduPIC data loaded: %zu electrons %zu ions, %.0f cycles completed before\n",
                    st.sp[0][0].size(), st.sp[1][0].size(), st.cycles);
    } else {
        // main.rs:1652-1662: N_INIT electrons and ions at random positions
        Xo rng{{0x9E3779B9u + seed, 0x7F4A7C15u, 0x85EBCA6Bu, 0xC2B2AE35u}};
        const double sigma = std::sqrt(1.38064852e-23 * p.t_gas / 6.63352090e-26);
        for (int sp = 0; sp < 2; ++sp) {
            std::vector<float> x(n_init), vx(n_init, 0.f), vy(n_init, 0.f), vz(n_init, 0.f);
            for (long k = 0; k < n_init; ++k) {
                x[k] = (float)(rng.uniform() * p.gap);
                if (thermal) {
                    vx[k] = (float)(rng.normal() * sigma);
                    vy[k] = (float)(rng.normal() * sigma);
                    vz[k] = (float)(rng.normal() * sigma);
                }
            }
            CK(pic_sim_upload(sim, sp, x.data(), vx.data(), vy.data(), vz.data(), (uint32_t)n_init));
        }
    }
    std::printf(">> cudaPIC: data uploaded to gpu\ninitialization time: %.3f s\n",
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_init).count());

    std::vector<uint32_t> hist_e, hist_i;
    {
        uint32_t ne = 0, ni = 0;
        CK(pic_sim_counts(sim, &ne, &ni));
        hist_e.push_back(ne);
        hist_i.push_back(ni);
    }
    std::printf(">> cudaPIC: running %ld cycles x%d steps (%s path)...\n", num_cycles, p.n_t, path.c_str());
    auto t0 = std::chrono::steady_clock::now();
    if (measure && xt_bin) CK(pic_sim_set_xt(sim, xt_bin));
    if (measure) CK(pic_sim_reset_measurements(sim));
    for (long cycle = 0; cycle < num_cycles; ++cycle) {
        const bool in_window = measure && cycle >= measurement_start;
        CK(pic_sim_run(sim, (uint32_t)p.n_t, in_window ? 1 : 0));
        if ((cycle + 1) % kCheckpointCycles == 0) {
            uint32_t ne = 0, ni = 0;
            CK(pic_sim_counts(sim, &ne, &ni));
            std::printf("   checkpoint at cycle %ld: n_e=%u, n_i=%u, time=%.3fs\n", cycle + 1, ne, ni,
                        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            hist_e.push_back(ne);
            hist_i.push_back(ni);
        }
    }
    CK(pic_sim_synchronize(sim));
    const int ovf = pic_sim_check(sim);
    if (ovf != PIC_OK)
        std::fprintf(stderr, ">> cudaPIC: WARNING: %s (raise --capacity)\n", pic_last_error());
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::vector<float> ex(capacity), evx(capacity), evy(capacity), evz(capacity);
    std::vector<float> ix(capacity), ivx(capacity), ivy(capacity), ivz(capacity);
    uint32_t ne = 0, ni = 0;
    CK(pic_sim_download(sim, PIC_ELECTRONS, ex.data(), evx.data(), evy.data(), evz.data(), (uint32_t)capacity, &ne));
    CK(pic_sim_download(sim, PIC_IONS, ix.data(), ivx.data(), ivy.data(), ivz.data(), (uint32_t)capacity, &ni));
    std::printf(">> cudaPIC: simulation complete in %.3f s\n", elapsed);
    std::printf(">> cudaPIC: final particles: %u electrons, %u ions\n", ne, ni);

    const std::string ts = timestamp();
    const std::string dir = out_root + "/" + ts;
    mkdirs(dir);
    auto write_particles = [&](const char *name, const std::vector<float> &x, const std::vector<float> &vx,
                               const std::vector<float> &vy, const std::vector<float> &vz, uint32_t n) {
        char fn[512];
        std::snprintf(fn, sizeof(fn), "%s/%04ld_%s_%s.csv", dir.c_str(), num_cycles, ts.c_str(), name);
        FILE *f = std::fopen(fn, "w");
        if (!f) return;
        std::fprintf(f, "x,vx,vy,vz\n");
        for (uint32_t k = 0; k < n && k < x.size(); ++k)
            std::fprintf(f, "%.9g,%.9g,%.9g,%.9g\n", x[k], vx[k], vy[k], vz[k]);
        std::fclose(f);
    };
    write_particles("electrons", ex, evx, evy, evz, ne);
    write_particles("ions", ix, ivx, ivy, ivz, ni);
    {
        FILE *f = std::fopen((dir + "/particle_growth_" + ts + ".csv").c_str(), "w");
        if (f) {
            std::fprintf(f, "step,n_e,n_i\n");
            for (size_t k = 0; k < hist_e.size(); ++k)
                std::fprintf(f, "%zu,%u,%u\n", k * kCheckpointCycles, hist_e[k], hist_i[k]);
            std::fclose(f);
        }
    }
    if (measure) {
        std::vector<double> ce(p.n_g), ci(p.n_g);
        std::vector<uint32_t> eepf(PIC_N_EEPF), ifed(2 * PIC_N_IFED);
        unsigned long long cnt[8];
        CK(pic_sim_read_measurements(sim, ce.data(), ci.data(), eepf.data(), ifed.data(), cnt));
        // main.rs:1867-1870: averages over the electron / ion steps measured
        const double n_steps_e = (double)cnt[6], n_steps_i = (double)cnt[7];
        FILE *f = std::fopen((dir + "/density_avg_" + ts + ".csv").c_str(), "w");
        if (f) {
            std::fprintf(f, "x,n_e,n_i\n");
            for (int k = 0; k < p.n_g; ++k)
                std::fprintf(f, "%.9g,%.9g,%.9g\n", k * p.dx, n_steps_e > 0 ? ce[k] / n_steps_e : 0.0,
                             n_steps_i > 0 ? ci[k] / n_steps_i : 0.0);
            std::fclose(f);
        }
        // main.rs:1592-1609: EEPF normalised by sum * dE * sqrt(E)
        double h = 0.0;
        for (uint32_t c : eepf) h += (double)c;
        h *= 0.05;
        f = std::fopen((dir + "/eepf_" + ts + ".csv").c_str(), "w");
        if (f) {
            std::fprintf(f, "energy_eV,eepf\n");
            for (int b = 0; b < PIC_N_EEPF; ++b) {
                const double e = (0.5 + b) * 0.05;
                std::fprintf(f, "%.6g,%.9g\n", e, h > 0 ? eepf[b] / h / std::sqrt(e) : 0.0);
            }
            std::fclose(f);
        }
        // C/eduPIC.cc:820-843: IFED normalised per electrode, mean ion energies
        double hp = 0.0, hg = 0.0, mp = 0.0, mg = 0.0;
        for (int b = 0; b < PIC_N_IFED; ++b) {
            hp += ifed[b];
            hg += ifed[PIC_N_IFED + b];
        }
        f = std::fopen((dir + "/ifed_" + ts + ".csv").c_str(), "w");
        if (f) {
            std::fprintf(f, "energy_eV,ifed_powered,ifed_grounded\n");
            for (int b = 0; b < PIC_N_IFED; ++b) {
                const double e = b + 0.5;
                const double fp = hp > 0 ? ifed[b] / hp : 0.0, fg = hg > 0 ? ifed[PIC_N_IFED + b] / hg : 0.0;
                mp += e * fp;
                mg += e * fg;
                std::fprintf(f, "%.2f,%.6f,%.6f\n", e, fp, fg);
            }
            std::fclose(f);
        }
        // XT maps first: the report's <j*E> power densities average them
        uint32_t n_xt = 0;
        std::vector<double> xt;
        if (xt_bin) {
            CK(pic_sim_read_xt(sim, nullptr, &n_xt));
            xt.resize((size_t)11 * p.n_g * n_xt);
            CK(pic_sim_read_xt(sim, xt.data(), &n_xt));
        }
        // C/eduPIC.cc:1002-1008: fluxes from the electrode absorption counters
        // over the measured time (the counters run over the whole run)
        const double t_meas = (double)measurement_cycles / p.freq;
        const double t_run = (double)num_cycles / p.freq;
        const double wa = p.weight / p.area;
        // C/eduPIC.cc:926-978: plasma characteristics and the stability and
        // accuracy conditions.  The mean electron energy in the EEPF window
        // comes from the EEPF counts at their bin centres (eduPIC sums the
        // energies themselves: at most DE_EEPF / 2 = 0.025 eV apart, and
        // electrons above the histogram's 100 eV are not in it).  An empty
        // window gives NaN, as eduPIC's 0/0, which passes its conditions.
        double cnt_c = 0.0, sum_c = 0.0;
        for (int b = 0; b < PIC_N_EEPF; ++b) {
            cnt_c += eepf[b];
            sum_c += eepf[b] * (b + 0.5) * 0.05;
        }
        pic_report_t rep{};
        {
            std::vector<float> cs((size_t)PIC_N_CS * PIC_CS_RANGES), se(PIC_CS_RANGES), si(PIC_CS_RANGES);
            CK(pic_cross_sections(p.gas_density, cs.data(), se.data(), si.data()));
            CK(pic_report(&p, se.data(), si.data(), n_steps_e > 0 ? ce[p.n_g / 2] / n_steps_e : 0.0,
                          cnt_c > 0 ? sum_c / cnt_c : std::nan(""), t_run, cnt[4], cnt[5], ne, ni, &rep));
        }
        f = std::fopen((dir + "/info_" + ts + ".txt").c_str(), "w");
        if (f) {
            const char *rule = "--------------------------------------------------------------------------------\n";
            std::fprintf(f, "cudaPIC (MI355X) run report\n");
            std::fprintf(f, "cycles run                            = %12ld\n", num_cycles);
            std::fprintf(f, "cycles measured                       = %12ld\n", measurement_cycles);
            std::fprintf(f, "%s", rule);
            std::fprintf(f, "Plasma characteristics:\n");
            std::fprintf(f, "Electron density @ center             = %12.3e [m^{-3}]\n", rep.density_center);
            std::fprintf(f, "Plasma frequency @ center             = %12.3e [rad/s]\n", rep.plasma_freq);
            std::fprintf(f, "Debye length @ center                 = %12.3e [m]\n", rep.debye_length);
            std::fprintf(f, "Electron collision frequency          = %12.3e [1/s]\n", rep.coll_freq_e);
            std::fprintf(f, "Ion collision frequency               = %12.3e [1/s]\n", rep.coll_freq_i);
            std::fprintf(f, "%s", rule);
            std::fprintf(f, "Stability and accuracy conditions:\n");
            std::fprintf(f, "Plasma frequency @ center * DT_E      = %12.3f (OK if less than 0.20)\n", rep.wp_dt_e);
            std::fprintf(f, "DX / Debye length @ center            = %12.3f (OK if less than 1.00)\n",
                         rep.dx_over_debye);
            std::fprintf(f, "Max. electron coll. frequency * DT_E  = %12.3f (OK if less than 0.05)\n",
                         rep.nu_max_e_dt_e);
            std::fprintf(f, "Max. ion coll. frequency * DT_I       = %12.3f (OK if less than 0.05)\n",
                         rep.nu_max_i_dt_i);
            std::fprintf(f, "%s", rule);
            if (!rep.conditions_ok) {
                std::fprintf(f, "** STABILITY AND ACCURACY CONDITION(S) VIOLATED - REFINE SIMULATION SETTINGS! **\n");
                std::fprintf(f, "%s", rule);
                std::printf(">> cudaPIC: ERROR: STABILITY AND ACCURACY CONDITION(S) VIOLATED (see info_%s.txt)\n",
                            ts.c_str());
            } else {
                std::fprintf(f, "Max e- energy for CFL condition       = %12.3f [eV]\n", rep.cfl_energy);
                std::fprintf(f, "%s", rule);
            }
            std::fprintf(f, "Particle characteristics at the electrodes:\n");
            std::fprintf(f, "Ion flux at powered electrode         = %12.3e [m^{-2} s^{-1}]\n", cnt[2] * wa / t_run);
            std::fprintf(f, "Ion flux at grounded electrode        = %12.3e [m^{-2} s^{-1}]\n", cnt[3] * wa / t_run);
            std::fprintf(f, "Mean ion energy at powered electrode  = %12.3e [eV]\n", mp);
            std::fprintf(f, "Mean ion energy at grounded electrode = %12.3e [eV]\n", mg);
            std::fprintf(f, "Electron flux at powered electrode    = %12.3e [m^{-2} s^{-1}]\n", cnt[0] * wa / t_run);
            std::fprintf(f, "Electron flux at grounded electrode   = %12.3e [m^{-2} s^{-1}]\n", cnt[1] * wa / t_run);
            std::fprintf(f, "Electron collisions                   = %12llu\n", cnt[4]);
            std::fprintf(f, "Ion collisions                        = %12llu\n", cnt[5]);
            std::fprintf(f, "measured time                         = %12.3e [s]\n", t_meas);
            if (xt_bin) {
                // C/eduPIC.cc:1012-1024: <j*E> averaged over the XT map
                double pe = 0.0, pi = 0.0;
                const size_t m = (size_t)p.n_g * n_xt;
                for (size_t k = 0; k < m; ++k) {
                    pe += xt[6 * m + k];
                    pi += xt[7 * m + k];
                }
                pe /= (double)m;
                pi /= (double)m;
                std::fprintf(f, "%s", rule);
                std::fprintf(f, "Absorbed power calculated as <j*E>:\n");
                std::fprintf(f, "Electron power density (average)      = %12.3e [W m^{-3}]\n", pe);
                std::fprintf(f, "Ion power density (average)           = %12.3e [W m^{-3}]\n", pi);
                std::fprintf(f, "Total power density(average)          = %12.3e [W m^{-3}]\n", pe + pi);
            }
            std::fclose(f);
        }
        if (xt_bin) {
            // save_all_xt (C/eduPIC.cc:905-919): N_G lines of N_XT values
            const char *names[11] = {"pot", "efield", "ne", "ni", "je", "ji",
                                     "powere", "poweri", "meanee", "meanei", "ioniz"};
            for (int q = 0; q < 11; ++q) {
                FILE *fx = std::fopen((dir + "/" + names[q] + "_xt.dat").c_str(), "w");
                if (!fx) continue;
                for (int k = 0; k < p.n_g; ++k) {
                    for (uint32_t b = 0; b < n_xt; ++b)
                        std::fprintf(fx, "%e  ", xt[((size_t)q * p.n_g + k) * n_xt + b]);
                    std::fprintf(fx, "\n");
                }
                std::fclose(fx);
            }
        }
        std::printf(">> cudaPIC: measurement data saved (density_avg, eepf, ifed, info%s) for %ld cycles\n",
                    xt_bin ? ", xt" : "", measurement_cycles);
    }
    if (!export_ed.empty()) {
        EduState st;
        st.time = ed_time + (double)num_cycles * p.n_t * p.dt_e;
        st.cycles = ed_cycles + (double)num_cycles;
        const std::vector<float> *src[2][4] = {{&ex, &evx, &evy, &evz}, {&ix, &ivx, &ivy, &ivz}};
        const uint32_t cnt[2] = {ne, ni};
        for (int sp = 0; sp < 2; ++sp)
            for (int c = 0; c < 4; ++c) st.sp[sp][c].assign(src[sp][c]->begin(), src[sp][c]->begin() + cnt[sp]);
        if (!write_edupic(export_ed, int_headers, st)) {
            std::fprintf(stderr, ">> cudaPIC: cannot write %s\n", export_ed.c_str());
            return 1;
        }
        std::printf(">> cudaPIC: eduPIC data saved to %s: %u electrons %u ions\n", export_ed.c_str(), ne, ni);
    }
    if (!save.empty()) {
        CK(pic_sim_save(sim, save.c_str()));
        std::printf(">> cudaPIC: state saved to %s\n", save.c_str());
    }
    std::printf(">> cudaPIC: results in %s\n", dir.c_str());
    pic_sim_destroy(sim);
    return ovf == PIC_OK ? 0 : 3;
}
