// pfsim -- host driver of the phase-field LBM timestep library.
//
// The reference's executable (src/main.cu:35-215), on the library's C-ABI:
//
//   pfsim <JET|DROPLET> <D3Q19|D3Q27> <ID> [options]
//
// Positional arguments, output directory bin/<CASE>/<VS>/<ID>/
// (host::createSimulationDirectory, src/functions/hostFunctions.cuh:33-46),
// the <ID>_info.txt metadata file and the console metadata block (:48-115),
// GPU_INDEX device selection (:117-150), the time loop STEP = 0..NSTEPS with
// output every MACRO_SAVE steps and at the last one (src/main.cu:163-186),
// and the closing execution-time and MLUPS report (:188-212) follow the
// reference.  What the reference fixes at compile time (flow case, velocity
// set, RUN/SAMPLE/PROFILE mode, BENCHMARK, grid, FP16 storage) is chosen here
// by the positional arguments and options, so one binary serves every case.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>

#include "phasefieldlbm.h"

namespace {

void usage(const char* argv0) {
    std::cerr << "Error: Usage: " << argv0 << " <flow case> <velocity set> <ID> [options]\n"
              << "  flow case     JET | DROPLET\n"
              << "  velocity set  D3Q19 | D3Q27\n"
              << "options:\n"
              << "  --mode run|sample|profile|benchmark  run length and output (default run:\n"
              << "                 100000 steps, output every 1000; sample: 10000/1000;\n"
              << "                 profile: 0/1; benchmark: run length, no output)\n"
              << "  --steps N      last step NSTEPS (the loop runs steps 0..N)\n"
              << "  --save N       output interval MACRO_SAVE\n"
              << "  --preset NAME  grid and physics: reference (default), droplet_d3q19_64,\n"
              << "                 jet_d3q19_128x128x512, jet_d3q27_256x256x1024, jet_d3q27_128x128x512\n"
              << "  --grid NX NY NZ, --diam D, --radius R, --sigma S\n"
              << "  --fp16         binary16 population storage\n"
              << "  --fp16-ne      binary16 storage of the non-equilibrium remainder\n"
              << "  --moments      FP32 storage of each cell's collision-output coefficients\n"
              << "  --noise gaussian|uniform|none, --noise-amp A, --seed N\n"
              << "  --closure FILE learned curvature closure weights (6017 little-endian float32)\n"
              << "  --derived LIST average,reynolds,vorticity: derived fields, appended to the output\n"
              << "                 (the reference's TIME_AVERAGE, REYNOLDS_MOMENTS, VORTICITY_FIELDS)\n"
              << "  --fields LIST  output fields, comma separated (default rho,uz)\n"
              << "  --format LIST  bin,vti,vtr (default bin,vtr)\n"
              << "  --root DIR     directory holding bin/ (default: current directory)\n"
              << "  --graph on|off replay each step's kernels as a captured hipGraph (default off)\n";
}

bool check(int rc, pf_ctx* ctx, const char* what) {
    if (rc == PF_OK) return true;
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, pf_status_string(rc), ctx ? pf_last_error(ctx) : "");
    return false;
}

std::vector<std::string> split(const std::string& s) {
    std::vector<std::string> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) out.push_back(item);
    return out;
}

int field_id(const std::string& name) {
    for (int32_t i = 0; i < PF_NUM_FIELDS; ++i)
        if (name == pf_field_name(i)) return i;
    return -1;
}

// Named workloads: the same physics at the validation and measurement grids.
bool apply_preset(pf_config& c, const std::string& name) {
    if (name == "reference") return true;
    if (name == "droplet_d3q19_64") {
        if (c.flow_case != PF_DROPLET) return false;
        c.nx = c.ny = c.nz = 64;
        c.radius = 12.0f;
        c.axis_x = c.axis_y = c.axis_z = 1.0f;
        c.sigma = 0.01f;
        c.tau = 0.55f;
        c.tau_g = 0.8f;
        return true;
    }
    if (c.flow_case != PF_JET) return false;
    c.noise = PF_NOISE_UNIFORM_Z;
    c.noise_amp = 0.05f;
    if (name == "jet_d3q19_128x128x512" || name == "jet_d3q27_128x128x512") {
        c.nx = 128, c.ny = 128, c.nz = 512, c.diam = 16;
        return true;
    }
    if (name == "jet_d3q27_256x256x1024") {
        c.nx = 256, c.ny = 256, c.nz = 1024, c.diam = 32;
        return true;
    }
    return false;
}

struct Run {
    std::string flow_case, velocity_set, id;
    long long nsteps = 100000, save = 1000;
    bool output = true;
    std::vector<int32_t> fields;
    int32_t formats = PF_OUT_BIN | PF_OUT_VTR;
    std::string closure_file, root;
    // Replay the step as a captured graph, as the reference does.  Off by
    // default: on MI355X a graph launch per step measured slower than direct
    // launches (64^3 droplet 24.6 against 18.1 us per step).
    bool graph = false;
};

void write_info(const std::string& dir, const Run& r, const pf_config& c, float sigma) {
    const std::string path = dir + r.id + "_info.txt";
    std::ofstream f(path, std::ios::out | std::ios::trunc);
    if (!f) {
        std::cerr << "Error opening file: " << path << std::endl;
        return;
    }
    const int diam = c.flow_case == PF_JET ? c.diam : (int)(2 * c.radius);
    f << "---------------------------- SIMULATION METADATA ----------------------------\n"
      << "ID:                    " << r.id << '\n'
      << "Velocity set:          " << r.velocity_set << '\n'
      << "Reference velocity:    " << (c.flow_case == PF_JET ? c.u_inf : 0.0f) << '\n'
      << "Dispersed phase Re:    " << (c.flow_case == PF_JET ? c.reynolds_oil : 0.0f) << '\n'
      << "Continuous phase Re:   " << (c.flow_case == PF_JET ? c.reynolds_water : 0.0f) << '\n'
      << "Weber number:          " << (c.flow_case == PF_JET ? c.weber : 0.0f) << '\n'
      << "Surface tension:       " << sigma << '\n'
      << "NX:                    " << c.nx << '\n'
      << "NY:                    " << c.ny << '\n'
      << "NZ:                    " << c.nz << '\n'
      << "Diameter:              " << diam << '\n'
      << "Timesteps:             " << r.nsteps << '\n'
      << "Output interval:       " << r.save << '\n'
      << "-----------------------------------------------------------------------------\n";
    std::cout << "Simulation information file created in: " << path << std::endl;
}

void print_metadata(const Run& r, const pf_config& c, float sigma) {
    // Mach = u * sqrt(3) (u / c_s); tau = 1/2 + 3 nu with nu = u D / Re.  A
    // droplet has no reference velocity: its single relaxation time is shown.
    const double as2 = 3.0;
    double tau_d = c.tau, tau_c = c.tau, ma = 0.0;
    if (c.flow_case == PF_JET) {
        ma = (double)c.u_inf * std::sqrt(as2);
        tau_d = 0.5 + (double)c.u_inf * c.diam / c.reynolds_oil * as2;
        tau_c = 0.5 + (double)c.u_inf * c.diam / c.reynolds_water * as2;
    }
    std::cout << "\n---------------------------- SIMULATION METADATA ----------------------------\n"
              << "Velocity set:          " << r.velocity_set << '\n'
              << "Reference velocity:    " << (c.flow_case == PF_JET ? c.u_inf : 0.0f) << '\n'
              << "Dispersed phase Re:    " << (c.flow_case == PF_JET ? c.reynolds_oil : 0.0f) << '\n'
              << "Continuous phase Re:   " << (c.flow_case == PF_JET ? c.reynolds_water : 0.0f) << '\n'
              << "Weber number:          " << (c.flow_case == PF_JET ? c.weber : 0.0f) << '\n'
              << "Surface tension:       " << sigma << '\n'
              << "NX:                    " << c.nx << '\n'
              << "NY:                    " << c.ny << '\n'
              << "NZ:                    " << c.nz << '\n'
              << "Diameter:              " << (c.flow_case == PF_JET ? c.diam : (int)(2 * c.radius)) << '\n'
              << "Total domain size:     " << (long long)c.nx * c.ny * c.nz << '\n'
              << "Mach:                  " << ma << '\n'
              << "Dispersed phase tau:   " << tau_d << '\n'
              << "Continuous phase tau:  " << tau_c << '\n'
              << "Storage:               " << (c.storage == PF_STORE_FP16 ? "FP16" : (c.storage == PF_STORE_FP16_NE ? "FP16 (non-equilibrium)" : (c.storage == PF_STORE_MOM ? "FP32 (moments)" : "FP32"))) << '\n'
              << "Curvature closure:     " << (c.closure == PF_CLOSURE_MLP ? "learned" : "finite difference")
              << '\n'
              << "-----------------------------------------------------------------------------\n\n";
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc < 4) {
        usage(argv[0]);
        return 1;
    }
    Run r;
    r.flow_case = argv[1];
    r.velocity_set = argv[2];
    r.id = argv[3];
    int32_t fc, vs;
    if (r.flow_case == "JET") fc = PF_JET;
    else if (r.flow_case == "DROPLET") fc = PF_DROPLET;
    else {
        std::cerr << "Error: unknown flow case '" << r.flow_case << "' (JET or DROPLET)\n";
        return 1;
    }
    if (r.velocity_set == "D3Q19") vs = PF_D3Q19;
    else if (r.velocity_set == "D3Q27") vs = PF_D3Q27;
    else {
        std::cerr << "Error: unknown velocity set '" << r.velocity_set << "' (D3Q19 or D3Q27)\n";
        return 1;
    }
    pf_config cfg;
    pf_config_default(&cfg, fc, vs);
    // The preset is applied first, whatever its position, so that the
    // explicit options refine it.
    std::string preset = "reference";
    for (int i = 4; i + 1 < argc; ++i)
        if (std::string(argv[i]) == "--preset") preset = argv[i + 1];
    if (!apply_preset(cfg, preset)) {
        std::cerr << "Error: preset '" << preset << "' does not exist for " << r.flow_case << "\n";
        return 1;
    }
    std::string fields = "rho,uz", formats = "bin,vtr";
    long long steps = -1, save = -1;
    std::string mode = "run";
    float sigma_override = -1.0f;
    for (int i = 4; i < argc; ++i) {
        const std::string a = argv[i];
        auto need = [&](int k) {
            if (i + k >= argc) {
                std::cerr << "Error: " << a << " needs " << k << " value(s)\n";
                std::exit(1);
            }
        };
        if (a == "--mode") { need(1); mode = argv[++i]; }
        else if (a == "--steps") { need(1); steps = std::atoll(argv[++i]); }
        else if (a == "--save") { need(1); save = std::atoll(argv[++i]); }
        else if (a == "--preset") { need(1); ++i; }
        else if (a == "--grid") { need(3); cfg.nx = std::atoi(argv[i + 1]); cfg.ny = std::atoi(argv[i + 2]); cfg.nz = std::atoi(argv[i + 3]); i += 3; }
        else if (a == "--diam") { need(1); cfg.diam = std::atoi(argv[++i]); }
        else if (a == "--radius") { need(1); cfg.radius = (float)std::atof(argv[++i]); }
        else if (a == "--sigma") { need(1); sigma_override = (float)std::atof(argv[++i]); }
        else if (a == "--fp16") cfg.storage = PF_STORE_FP16;
        else if (a == "--fp16-ne") cfg.storage = PF_STORE_FP16_NE;
        else if (a == "--moments") cfg.storage = PF_STORE_MOM;
        else if (a == "--noise") {
            need(1);
            const std::string n = argv[++i];
            if (n == "gaussian") cfg.noise = PF_NOISE_GAUSSIAN_XY;
            else if (n == "uniform") cfg.noise = PF_NOISE_UNIFORM_Z;
            else if (n == "none") cfg.noise = PF_NOISE_NONE;
            else { std::cerr << "Error: unknown noise '" << n << "'\n"; return 1; }
        }
        else if (a == "--noise-amp") { need(1); cfg.noise_amp = (float)std::atof(argv[++i]); }
        else if (a == "--seed") { need(1); cfg.seed = (uint32_t)std::strtoul(argv[++i], nullptr, 0); }
        else if (a == "--closure") { need(1); r.closure_file = argv[++i]; }
        else if (a == "--graph") {
            need(1);
            const std::string v = argv[++i];
            if (v == "on") r.graph = true;
            else if (v == "off") r.graph = false;
            else { std::cerr << "Error: --graph takes on or off\n"; return 1; }
        }
        else if (a == "--fields") { need(1); fields = argv[++i]; }
        else if (a == "--derived") {
            need(1);
            for (const std::string& d : split(argv[++i])) {
                if (d == "average") cfg.derived |= PF_DERIVED_AVERAGE;
                else if (d == "reynolds") cfg.derived |= PF_DERIVED_REYNOLDS;
                else if (d == "vorticity") cfg.derived |= PF_DERIVED_VORTICITY;
                else { std::cerr << "Error: unknown derived field set '" << d << "'\n"; return 1; }
            }
        }
        else if (a == "--format") { need(1); formats = argv[++i]; }
        else if (a == "--root") { need(1); r.root = argv[++i]; }
        else { std::cerr << "Error: unknown option '" << a << "'\n"; usage(argv[0]); return 1; }
    }
    if (sigma_override >= 0.0f) cfg.sigma = sigma_override;
    if (!r.closure_file.empty()) cfg.closure = PF_CLOSURE_MLP;

    if (mode == "run") r.nsteps = 100000, r.save = 1000;
    else if (mode == "sample") r.nsteps = 10000, r.save = 1000;
    else if (mode == "profile") r.nsteps = 0, r.save = 1;
    else if (mode == "benchmark") r.nsteps = 100000, r.save = 1000, r.output = false;
    else { std::cerr << "Error: unknown mode '" << mode << "'\n"; return 1; }
    if (steps >= 0) r.nsteps = steps;
    if (save >= 0) r.save = save;
    if (r.save < 1) { std::cerr << "Error: --save must be at least 1\n"; return 1; }
    for (const std::string& f : split(fields)) {
        const int id = field_id(f);
        if (id < 0) { std::cerr << "Error: unknown field '" << f << "'\n"; return 1; }
        r.fields.push_back(id);
    }
    // enabled derived fields follow the base fields (src/main.cu:130-146)
    for (int32_t id = PF_NUM_BASE_FIELDS; id < PF_NUM_FIELDS; ++id) {
        const bool on = (id <= PF_FIELD_AVG_UZ && (cfg.derived & PF_DERIVED_AVERAGE)) ||
                        (id >= PF_FIELD_AVG_UXUX && id <= PF_FIELD_AVG_UYUZ && (cfg.derived & PF_DERIVED_REYNOLDS)) ||
                        (id >= PF_FIELD_VORT_X && (cfg.derived & PF_DERIVED_VORTICITY));
        bool listed = false;
        for (int32_t f : r.fields) listed = listed || f == id;
        if (on && !listed) r.fields.push_back(id);
    }
    r.formats = 0;
    for (const std::string& f : split(formats)) {
        if (f == "bin") r.formats |= PF_OUT_BIN;
        else if (f == "vti") r.formats |= PF_OUT_VTI;
        else if (f == "vtr") r.formats |= PF_OUT_VTR;
        else { std::cerr << "Error: unknown format '" << f << "'\n"; return 1; }
    }
    if (r.output && (r.fields.empty() || r.formats == 0)) { std::cerr << "Error: nothing to write\n"; return 1; }

    namespace fs = std::filesystem;
    const fs::path base = r.root.empty() ? fs::current_path() : fs::path(r.root);
    const fs::path sim_dir = base / "bin" / r.flow_case / r.velocity_set / r.id;
    std::error_code ec;
    fs::create_directories(sim_dir, ec);
    const std::string SIM_DIR = sim_dir.string() + "/";

    pf_ctx* ctx = nullptr;
    int rc = pf_create(&cfg, &ctx);
    if (rc != PF_OK) {
        std::fprintf(stderr, "pf_create failed: %s\n", pf_status_string(rc));
        return 1;
    }
    pf_get_config(ctx, &cfg);
    char name[256] = "unknown";
    pf_device_name(cfg.device, name, sizeof name);
    std::printf("Using GPU %d: %s\n", cfg.device, name);
    std::fflush(stdout);

    if (cfg.closure == PF_CLOSURE_MLP) {
        std::vector<float> p(PF_MLP_PARAMS);
        std::ifstream in(r.closure_file, std::ios::binary);
        if (!in.read(reinterpret_cast<char*>(p.data()), (std::streamsize)(p.size() * sizeof(float)))) {
            std::cerr << "Error: cannot read " << PF_MLP_PARAMS << " float32 values from " << r.closure_file << "\n";
            pf_destroy(ctx);
            return 1;
        }
        const float* q = p.data();
        if (!check(pf_set_closure(ctx, q, q + 1728, q + 1792, q + 5888, q + 5952, q + 6016), ctx, "pf_set_closure")) {
            pf_destroy(ctx);
            return 1;
        }
    }
    if (!check(pf_init(ctx), ctx, "pf_init") || !check(pf_sync(ctx), ctx, "pf_sync")) {
        pf_destroy(ctx);
        return 1;
    }
    // The jet's surface tension is derived (u^2 D / We); report the value in use.
    const float sigma = cfg.flow_case == PF_JET ? cfg.u_inf * cfg.u_inf * (float)cfg.diam / cfg.weber : cfg.sigma;
    write_info(SIM_DIR, r, cfg, sigma);
    print_metadata(r, cfg, sigma);

    const auto t_start = std::chrono::high_resolution_clock::now();
    long long step = 0;
    bool ok = true;
    while (ok && step <= r.nsteps) {
        // next output step: a multiple of MACRO_SAVE, or the last step
        long long next = r.nsteps;
        if (r.output) {
            const long long m = ((step + r.save - 1) / r.save) * r.save;
            next = std::min(m, r.nsteps);
        }
        const long long n = next - step + 1;
        for (long long done = 0; ok && done < n;) {
            const uint32_t chunk = (uint32_t)std::min<long long>(n - done, 1 << 20);
            const bool last = done + chunk == n;
            const uint32_t flags = ((r.output && last) ? PF_STEP_SAVE_LAST : 0u) | (r.graph ? PF_STEP_GRAPH : 0u);
            ok = check(pf_step(ctx, (uint64_t)(step + done), chunk, flags), ctx, "pf_step");
            done += chunk;
        }
        if (ok && r.output) {
            ok = check(pf_sync(ctx), ctx, "pf_sync");
            if (ok) {
                std::cout << "Step " << next << ": bins in " << SIM_DIR << "\n" << std::flush;
                ok = check(pf_write_fields(ctx, SIM_DIR.c_str(), (uint64_t)next, r.fields.data(),
                                           (int32_t)r.fields.size(), r.formats),
                           ctx, "pf_write_fields");
            }
        }
        step = next + 1;
    }
    ok = ok && check(pf_sync(ctx), ctx, "pf_sync");
    const auto t_end = std::chrono::high_resolution_clock::now();
    pf_destroy(ctx);
    if (!ok) return 1;

    const double elapsed = std::chrono::duration<double>(t_end - t_start).count();
    const double updates = (double)cfg.nx * cfg.ny * cfg.nz * (double)(r.nsteps + 1);
    std::cout << "\n// =============================================== //\n";
    std::cout << "     Total execution time    : " << elapsed << " s\n";
    std::cout << "     Performance             : " << updates / (elapsed * 1e6) << " MLUPS\n";
    std::cout << "// =============================================== //\n\n";
    return 0;
}
