// eu3d_run: the reference's driver flow (Euler constructor + Computing) from its control deck,
// on the GPU through the host layer. The flow is eu3d_run_main in libeu3d_host.so, callable
// in-process with a command line; the eu3d_run program (eu3d_run_main.cpp) only forwards to it.
//
//   eu3d_run <ctrl deck> [--ic <file>] [--bc k0,k1,k2,k3,k4,k5] [--steps N] [--fused RK]
//            [--quirks ref|fixed] [--out DIR] [--dump FILE] [--time-record]
//   eu3d_run <ctrl deck> --print-decks     (parse only, no GPU: JSON of what the decks yield)
//
// The control deck names the grid deck, the mechanism and the thermo file as in the reference.
// --ic gives the primitive fields in reference layout (ghosts included) as raw float64: T, P, U,
// V, W (one value per cell each) then Yi (NS values per cell). Without --ic the run starts from
// the reference's own initial state: the shock deck that the control deck's "initialization"
// names, ReactionInitial and FieldInitial (the 3dRSBI start). --bc replaces
// FieldBoundary_3dRSBI with the given ghost kinds. --fused runs the steps as one eu3d_step call
// with RK stages (1 or 2) instead of the method-by-method loop. --dump writes CS (13 per cell)
// and T after the run as raw float64 in reference layout. --time-record times the phases of the
// method-by-method loop and writes the reference's time-record files under DIR/output_1/.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

#include "eu3d_host.hpp"

using namespace eu3d_host;

static void json_array(const char* name, const std::vector<double>& v, bool last = false) {
    std::printf("\"%s\": [", name);
    for (size_t i = 0; i < v.size(); i++) std::printf("%s%.17g", i ? ", " : "", v[i]);
    std::printf("]%s\n", last ? "" : ",");
}

// Control, grid and mechanism as the decks give them, as one JSON object on stdout.
static int print_decks(const std::string& ctrl) {
    Control c;
    c.FileRead(ctrl);
    FileReader raw;
    raw.readFile(ctrl);
    Mesh m;
    m.MeshProcess(c.gridname);
    Mechanism me;
    me.ReactionRead(c.reaction_model, c.Thermofile);
    std::printf("{\"deck\": {");
    bool first = true;
    for (const auto& kv : raw.all()) {
        std::printf("%s\"%s\": \"%s\"", first ? "" : ", ", kv.first.c_str(), kv.second.c_str());
        first = false;
    }
    std::printf("},\n\"control\": {\"Final_Time\": %.17g, \"cfl\": %.17g, \"dt\": %.17g, \"count\": %d, "
                "\"Reaction_Sch\": %d, \"type\": %d, \"Diff_Sch\": %d, \"TimeAdv_Sch\": %d},\n",
                c.Final_Time, c.cfl, c.dt, c.count, c.Reaction_Sch, c.type, c.Diff_Sch, c.TimeAdv_Sch);
    std::printf("\"mesh\": {\"ni\": %d, \"nj\": %d, \"nk\": %d, \"bc\": %d,\n", m.ni, m.nj, m.nk, m.bc);
    json_array("xnode", m.xnode);
    json_array("ynode", m.ynode);
    json_array("znode", m.znode, true);
    std::printf("},\n\"mech\": {\"NS\": %d, \"NR\": %d, \"TB\": %d, \"species\": [", me.NS, me.NR, me.TB);
    for (size_t i = 0; i < me.species.size(); i++) std::printf("%s\"%s\"", i ? ", " : "", me.species[i].c_str());
    std::printf("],\n");
    json_array("stoi_f", me.stoi_f);
    json_array("stoi_b", me.stoi_b);
    json_array("af", me.af);
    json_array("bf", me.bf);
    json_array("eaf", me.eaf);
    json_array("react_tb", me.react_tb);
    json_array("coeff0", me.coeff0);
    json_array("coeff1", me.coeff1);
    json_array("mw", me.mw);
    json_array("mr", me.mr, true);
    std::printf("}}\n");
    std::fflush(stdout);
    return 0;
}

static std::vector<double> read_block(std::ifstream& in, size_t n, const char* what) {
    std::vector<double> v(n);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(double)));
    if (!in) throw Error(EU3D_ERR_ARG, std::string("initial-condition file too short at ") + what);
    return v;
}

extern "C" int eu3d_run_main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <ctrl deck> [--ic FILE] [--bc k,k,k,k,k,k] [--steps N] [--fused RK] "
                             "[--quirks ref|fixed] [--out DIR] [--dump FILE]\n", argv[0]);
        return 2;
    }
    std::string ctrl = argv[1], ic, out = ".", dump;
    if (argc == 3 && std::string(argv[2]) == "--print-decks") {
        try {
            return print_decks(ctrl);
        } catch (const Error& err) {
            std::fprintf(stderr, "eu3d_run: %s (status %d)\n", err.what(), err.code);
            return 1;
        }
    }
    int steps = 0, fused = 0;
    bool have_bc = false, time_record = false;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    uint32_t quirks = EU3D_QUIRKS_FIXED;
    for (int a = 2; a < argc; a++) {
        const std::string k = argv[a];
        auto val = [&]() -> std::string {
            if (a + 1 >= argc) throw Error(EU3D_ERR_ARG, "missing value for " + k);
            return argv[++a];
        };
        if (k == "--ic") ic = val();
        else if (k == "--steps") steps = std::atoi(val().c_str());
        else if (k == "--fused") fused = std::atoi(val().c_str());
        else if (k == "--out") out = val();
        else if (k == "--dump") dump = val();
        else if (k == "--quirks") quirks = val() == "ref" ? EU3D_QUIRKS_REF : EU3D_QUIRKS_FIXED;
        else if (k == "--time-record") time_record = true;
        else if (k == "--bc") {
            std::stringstream ss(val());
            std::string t;
            for (int f = 0; f < 6 && std::getline(ss, t, ','); f++) bc[f] = std::atoi(t.c_str());
            have_bc = true;
        } else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            return 2;
        }
    }
    try {
        auto clock = []() {
            return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
        };
        const double t0 = clock();
        Euler e(ctrl, quirks);
        e.set_output_dir(out);
        e.set_time_record(time_record);
        const double t1 = clock();
        const size_t n3 = e.mesh.cells_with_ghosts();
        double t2;
        if (ic.empty()) {
            // the reference's own start: the shock deck, ReactionInitial, FieldInitial
            t2 = clock();
            e.Initialize();
        } else {
            std::ifstream in(ic, std::ios::binary);
            if (!in.is_open()) throw Error(EU3D_ERR_ARG, "cannot read initial condition " + ic);
            std::vector<double> T = read_block(in, n3, "T"), P = read_block(in, n3, "P"),
                                U = read_block(in, n3, "U"), V = read_block(in, n3, "V"),
                                W = read_block(in, n3, "W"), Y = read_block(in, n3 * e.mech.NS, "Yi");
            t2 = clock();
            e.TwoDim.SetPrimitives(T, P, U, V, W, Y);
            e.TwoDim.FieldInitial();
        }
        e.TwoDim.IsNan();   // waits for the device
        const double t3 = clock();
        if (fused) {
            if (!have_bc) throw Error(EU3D_ERR_ARG, "--fused needs --bc");
            e.ComputingFused(steps, bc, fused);
        } else {
            e.Computing(e.ctrl.count, have_bc ? bc : nullptr, steps);
        }
        e.TwoDim.IsNan();
        const double t4 = clock();
        if (time_record) {
            // the decks, grid and mechanism are read together by the constructor: FileReader
            // holds that time, MeshProcess and ReactionRead 0
            e.Output_Total_Time(t1 - t0, 0.0, t2 - t1, 0.0, t3 - t2, t4 - t3, 0.0);
            e.Output_Compute_Time();
        }
        std::printf("iterations %d time %.9e dt %.9e\n", e.iteration, e.TwoDim.time(), e.TwoDim.dt());
        std::fflush(stdout);
        if (!dump.empty()) {
            std::ofstream d(dump, std::ios::binary);
            const std::vector<double> cs = e.TwoDim.field(EU3D_FIELD_CS), t = e.TwoDim.field(EU3D_FIELD_T);
            d.write(reinterpret_cast<const char*>(cs.data()), (std::streamsize)(cs.size() * sizeof(double)));
            d.write(reinterpret_cast<const char*>(t.data()), (std::streamsize)(t.size() * sizeof(double)));
        }
    } catch (const Error& err) {
        std::fprintf(stderr, "eu3d_run: %s (status %d)\n", err.what(), err.code);
        return 1;
    }
    return 0;
}
