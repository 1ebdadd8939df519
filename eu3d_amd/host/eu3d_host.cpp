// Implementation of the C++ host layer (see eu3d_host.hpp). Plain C++17, no HIP: every device
// operation goes through the C ABI of libeu3d.so.
#include "eu3d_host.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>

namespace eu3d_host {

// ------------------------------------------------------------------ FileReader
namespace {
// std::stoi / std::stod success test without exceptions escaping
bool leading_int(const std::string& s) {
    try {
        (void)std::stoi(s);
        return true;
    } catch (...) {
        return false;
    }
}
std::string six_digits(double v) {
    std::ostringstream ss;   // default stream precision: 6 significant digits
    ss << v;
    return ss.str();
}
// create a directory and its parents; failure is an Error like the other output failures
void make_dirs(const std::string& dir) {
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) throw Error(EU3D_ERR_ARG, "cannot create " + dir + ": " + ec.message());
}
}  // namespace

bool FileReader::readFile(const std::string& name) {
    std::ifstream in(name);
    if (!in.is_open()) return false;
    std::string line, key, value;   // key / value persist across lines, as extraction leaves them
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        ss >> key >> value;
        if (key.empty() || value.empty() || key[0] == '#' || value[0] == '#') continue;
        try {
            const double v = std::stod(value);
            parameters[key] = six_digits(v);
        } catch (...) {
            parameters[key] = value;
        }
    }
    return true;
}

const std::string& FileReader::at(const std::string& key) const {
    auto it = parameters.find(key);
    if (it == parameters.end()) throw Error(EU3D_ERR_ARG, "deck key not found: " + key);
    return it->second;
}

int FileReader::getIntParameter(const std::string& key) const {
    const std::string& v = at(key);
    return leading_int(v) ? std::stoi(v) : 0;
}

double FileReader::getdoubleParameter(const std::string& key) const {
    const std::string& v = at(key);
    return leading_int(v) ? std::stod(v) : 0.0;
}

std::string FileReader::getStringParameter(const std::string& key) const { return at(key); }

bool FileReader::findparameters(const std::string& key) const { return parameters.count(key) != 0; }

void FileReader::setParameter(const std::string& key, const std::string& value) { parameters[key] = value; }

void FileReader::setParameter(const std::string& key, double value) { parameters[key] = six_digits(value); }

// ------------------------------------------------------------------ Mesh
void Mesh::Initial(int total_ni, double xa, double xb, int total_nj, double ya, double yb, int total_nk, double za,
                   double zb, int bc_) {
    bc = bc_;
    auto axis = [&](int total, double a, double b, int& n, double& d, std::vector<double>& node) {
        d = (b - a) / (total - 1);
        n = total;   // one block holds every node
        node.assign(n + 2 * bc, 0.0);
        for (int i = bc; i < n + bc; i++) node[i] = a + (i - bc) * d;
        // ghost nodes mirrored through the first / last interior node. An axis with fewer than
        // bc + 1 nodes (the 1-D decks' total_nj = 2) has no node to mirror: the reference then
        // reads a ghost entry not yet written (src/Mesh.cpp:114-118), which makes the CFL
        // spacing y[1] - y[0] negative; those ghosts extend the spacing d linearly instead.
        for (int i = 0; i < bc; i++) {
            node[i] = (2 * bc - i < n + bc) ? 2 * node[bc] - node[2 * bc - i] : node[bc] - (bc - i) * d;
            node[n + bc + i] = (n + bc - 2 - i >= bc) ? 2 * node[n + bc - 1] - node[n + bc - 2 - i]
                                                      : node[n + bc - 1] + (i + 1) * d;
        }
    };
    axis(total_ni, xa, xb, ni, dx, xnode);
    axis(total_nj, ya, yb, nj, dy, ynode);
    axis(total_nk, za, zb, nk, dz, znode);
}

void Mesh::MeshProcess(const std::string& gridfile) {
    FileReader g;
    if (!g.readFile(gridfile)) throw Error(EU3D_ERR_ARG, "cannot read grid deck " + gridfile);
    Initial(g.getIntParameter("total_ni"), g.getdoubleParameter("xa"), g.getdoubleParameter("xb"),
            g.getIntParameter("total_nj"), g.getdoubleParameter("ya"), g.getdoubleParameter("yb"),
            g.getIntParameter("total_nk"), g.getdoubleParameter("za"), g.getdoubleParameter("zb"),
            g.getIntParameter("bc"));
}

void Mesh::Uniform(int ni_, int nj_, int nk_, double d, int bc_) {
    Initial(ni_, 0.0, (ni_ - 1) * d, nj_, 0.0, (nj_ - 1) * d, nk_, 0.0, (nk_ - 1) * d, bc_);
    dx = dy = dz = d;
}

// ------------------------------------------------------------------ Mechanism
namespace {
std::vector<std::string> read_lines(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw Error(EU3D_ERR_ARG, "cannot read " + path);
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);   // '\r' stays, as getline leaves it
    return lines;
}

std::vector<std::string> split_tab(const std::string& s) {
    std::vector<std::string> out;
    std::string cur;
    std::istringstream ss(s);
    while (std::getline(ss, cur, '\t')) out.push_back(cur);
    return out;
}

// whitespace tokens from line `start` onwards
class Tokens {
  public:
    Tokens(const std::vector<std::string>& L, size_t start) : L(L), line(start) {}
    std::string next() {
        while (true) {
            std::string t;
            if (ss >> t) return t;
            if (line >= L.size()) throw Error(EU3D_ERR_ARG, "mechanism file ends early");
            ss.clear();
            ss.str(L[line++]);
        }
    }
    double next_double() {   // stream extraction of one token into double
        std::istringstream t(next());
        double v = 0;
        t >> v;
        return v;
    }
    int next_int() {         // stream extraction into int: "16." gives 16
        std::istringstream t(next());
        int v = 0;
        t >> v;
        return v;
    }

  private:
    const std::vector<std::string>& L;
    size_t line;
    std::istringstream ss;
};
}  // namespace

// Line layout of the mechanism file (src/Reaction.cpp:150-283): line 2 holds NS, NR, TB
// (tab separated); line 4 the species names; from line 8 two tab-separated stoichiometry rows
// per reaction (forward, backward); 2 lines further the Arrhenius rows "idx A b E"; 3 lines
// after those the third-body rows "reaction alpha_1 .. alpha_NS" read as integers; 2 lines
// further the initial molar ratios. Thermo file (:285-321): species j's 19 numbers (low range
// a1..a7 b1 b2, high range, Mw) start on line 15 + 6 j.
void Mechanism::ReactionRead(const std::string& reaction_model, const std::string& thermofile) {
    const std::vector<std::string> L = read_lines(reaction_model);
    if (L.size() < 9) throw Error(EU3D_ERR_ARG, "mechanism file too short");
    std::vector<int> head;
    for (const std::string& t : split_tab(L[2]))
        if (!t.empty()) head.push_back(std::stoi(t));
    if (head.size() < 3) throw Error(EU3D_ERR_ARG, "mechanism size line");
    NS = head[0], NR = head[1], TB = head[2];
    species.clear();
    for (std::string t : split_tab(L[4])) {
        while (!t.empty() && (t.back() == '\r' || t.back() == ' ')) t.pop_back();
        if (!t.empty() && (int)species.size() < NS) species.push_back(t);
    }
    stoi_f.assign((size_t)NS * NR, 0.0);
    stoi_b.assign((size_t)NS * NR, 0.0);
    size_t line = 8;
    for (int r = 0; r < NR; r++) {
        for (int side = 0; side < 2; side++) {
            if (line + 2 * r + side >= L.size()) throw Error(EU3D_ERR_ARG, "stoichiometry rows");
            const std::vector<std::string> row = split_tab(L[line + 2 * r + side]);
            for (int s = 0; s < NS && s < (int)row.size(); s++)
                (side ? stoi_b : stoi_f)[(size_t)s * NR + r] = std::stoi(row[s]);
        }
    }
    line += 2 * NR + 2;
    af.assign(NR, 0.0), bf.assign(NR, 0.0), eaf.assign(NR, 0.0);
    {
        Tokens t(L, line);
        for (int r = 0; r < NR; r++) {
            (void)t.next_double();
            af[r] = t.next_double();
            bf[r] = t.next_double();
            eaf[r] = t.next_double();
        }
    }
    line += NR + 3;
    react_tb.assign((size_t)NS * NR, 0.0);
    {
        Tokens t(L, line);
        for (int k = 0; k < TB; k++) {
            const int r = t.next_int() - 1;
            for (int s = 0; s < NS; s++) {
                const int a = t.next_int();
                if (r >= 0 && r < NR) react_tb[(size_t)s * NR + r] = a;
            }
        }
    }
    line += TB + 2;
    mr.assign(NS, 0.0);
    {
        Tokens t(L, line);
        for (int s = 0; s < NS; s++) mr[s] = t.next_double();
    }
    const std::vector<std::string> T = read_lines(thermofile);
    coeff0.assign((size_t)9 * NS, 0.0), coeff1.assign((size_t)9 * NS, 0.0), mw.assign(NS, 0.0);
    for (int s = 0; s < NS; s++) {
        Tokens t(T, 15 + 6 * (size_t)s);
        for (int k = 0; k < 9; k++) coeff0[(size_t)k * NS + s] = t.next_double();
        for (int k = 0; k < 9; k++) coeff1[(size_t)k * NS + s] = t.next_double();
        mw[s] = t.next_double();
    }
}

eu3d_mech Mechanism::view() const {
    eu3d_mech m;
    m.ns = NS;
    m.nr = NR;
    m.stoi_f = stoi_f.data();
    m.stoi_b = stoi_b.data();
    m.af = af.data();
    m.bf = bf.data();
    m.eaf = eaf.data();
    m.react_tb = react_tb.data();
    m.coeff0 = coeff0.data();
    m.coeff1 = coeff1.data();
    m.mw = mw.data();
    return m;
}

// ------------------------------------------------------------------ Context
Context::Context(const Mesh& mesh, const Mechanism& mech, uint32_t quirks) : m(mesh) {
    eu3d_grid g;
    g.ni = m.ni, g.nj = m.nj, g.nk = m.nk, g.bc = m.bc;
    g.xnode = m.xnode.data(), g.ynode = m.ynode.data(), g.znode = m.znode.data();
    const eu3d_mech mv = mech.view();
    const int rc = eu3d_create(&ctx, &g, &mv, quirks);
    if (rc != EU3D_OK) throw Error(rc, "eu3d_create failed (status " + std::to_string(rc) + ")");
}

Context::~Context() {
    if (ctx) eu3d_destroy(ctx);
}

void Context::check(int rc, const char* what) const {
    if (rc != EU3D_OK) {
        const char* msg = eu3d_last_error(ctx);
        throw Error(rc, std::string(what) + ": " + (msg ? msg : ""));
    }
}

// ------------------------------------------------------------------ Flowfield
void Flowfield::SetPrimitives(const std::vector<double>& T, const std::vector<double>& P, const std::vector<double>& U,
                              const std::vector<double>& V, const std::vector<double>& W, const std::vector<double>& Yi) {
    set_field(EU3D_FIELD_T, T);
    set_field(EU3D_FIELD_P, P);
    set_field(EU3D_FIELD_U, U);
    set_field(EU3D_FIELD_V, V);
    set_field(EU3D_FIELD_W, W);
    set_field(EU3D_FIELD_YI, Yi);
}

void Flowfield::FieldInitial() { ctx.check(eu3d_field_initial(ctx.get()), "FieldInitial"); }

void Flowfield::FieldInitial(const InitialDeck& in, const Mesh& m, const std::vector<double>& Yi) {
    const double M = in.M_wave, T_pre = in.T_pre, P_pre = in.P_pre;
    // the moving shock's post-shock state as src/Flowfield.cpp:172-174 forms it (ideal gas,
    // gamma 1.4, R 287 ahead of the shock)
    const double U_post = M * (1 - (2 + 0.40 * std::pow(M, 2)) / (2.40 * std::pow(M, 2))) * std::sqrt(1.40 * 287 * T_pre);
    const double P_post = P_pre * (1 + 2 * 1.4 / 2.4 * (std::pow(M, 2) - 1.0));
    const double T_post = T_pre * (2 + 0.40 * std::pow(M, 2)) / (2.40 * std::pow(M, 2)) * P_post / P_pre;
    const size_t NX = m.ni + 2 * m.bc, NYZ = (size_t)(m.nj + 2 * m.bc) * (m.nk + 2 * m.bc), n3 = NX * NYZ;
    std::vector<double> T(n3), P(n3), U(n3), zero(n3, 0.0);
    for (size_t i = 0; i < NX; i++) {
        const bool pre = m.xnode[i] >= 0.01;
        for (size_t c = i * NYZ; c < (i + 1) * NYZ; c++) {
            P[c] = pre ? P_pre : P_post;
            T[c] = pre ? T_pre : T_post;
            U[c] = pre ? 0.0 : U_post;
        }
    }
    SetPrimitives(T, P, U, zero, zero, Yi);
    FieldInitial();
}

void Flowfield::CFLcondition(double cfl, double Final_Time) {
    ctx.check(eu3d_cfl_dt(ctx.get(), cfl, Final_Time), "CFLcondition");
}

void Flowfield::FieldBoundary_3dRSBI() {
    const int k[6] = {EU3D_BC_RSBI, EU3D_BC_RSBI, EU3D_BC_RSBI, EU3D_BC_RSBI, EU3D_BC_RSBI, EU3D_BC_RSBI};
    FieldBoundary(k);
}

void Flowfield::FieldBoundary(const int kinds[6]) { ctx.check(eu3d_fill_ghosts(ctx.get(), kinds), "FieldBoundary"); }

void Flowfield::Advection(int TimeAdv_Sch, int Diff_Sch, int limit, bool fused) {
    if (TimeAdv_Sch != TimeAdv::EE) throw Error(EU3D_ERR_UNSUPPORTED, "Advection: only EE advances the state");
    const int rc = fused ? eu3d_advection_fused(ctx.get(), Diff_Sch, limit) : eu3d_advection(ctx.get(), Diff_Sch, limit);
    ctx.check(rc, "Advection");
}

void Flowfield::Update_after_Adv() { ctx.check(eu3d_update_after_adv(ctx.get()), "Update_after_Adv"); }
void Flowfield::Explicit() { ctx.check(eu3d_explicit(ctx.get()), "Explicit"); }
void Flowfield::Update_IMEX(Reaction&) { ctx.check(eu3d_update_imex(ctx.get()), "Update_IMEX"); }
void Flowfield::Update_after_CS() { ctx.check(eu3d_update_after_cs(ctx.get()), "Update_after_CS"); }
void Flowfield::GetPartial_T() { ctx.check(eu3d_get_partial_t(ctx.get()), "GetPartial_T"); }

void Flowfield::set_dt(double dt) {
    double d, t;
    ctx.check(eu3d_get_time(ctx.get(), &d, &t), "get_time");
    ctx.check(eu3d_set_time(ctx.get(), dt, t), "set_time");
}

double Flowfield::dt() const {
    double d, t;
    ctx.check(eu3d_get_time(ctx.get(), &d, &t), "get_time");
    return d;
}

double Flowfield::time() const {
    double d, t;
    ctx.check(eu3d_get_time(ctx.get(), &d, &t), "get_time");
    return t;
}

std::vector<double> Flowfield::field(int id) const {
    const int nc = eu3d_field_components(ctx.get(), id);
    if (nc <= 0) throw Error(EU3D_ERR_ARG, "unknown field id");
    std::vector<double> h(ctx.mesh().cells_with_ghosts() * nc);
    ctx.check(eu3d_get_field(ctx.get(), id, h.data()), "get_field");
    return h;
}

void Flowfield::set_field(int id, const std::vector<double>& host) {
    const int nc = eu3d_field_components(ctx.get(), id);
    if (nc <= 0 || host.size() != ctx.mesh().cells_with_ghosts() * nc)
        throw Error(EU3D_ERR_ARG, "set_field: size does not match the field");
    ctx.check(eu3d_set_field(ctx.get(), id, host.data()), "set_field");
}

bool Flowfield::IsNan() const {
    const int rc = eu3d_check_finite(ctx.get());
    if (rc == EU3D_ERR_NONFINITE) return true;
    ctx.check(rc, "check_finite");
    return false;
}

// ------------------------------------------------------------------ initial state decks
void InitialDeck::InputRead(const std::string& initialization) {
    FileReader in;
    if (!in.readFile(initialization)) throw Error(EU3D_ERR_ARG, "cannot read initialization deck " + initialization);
    P_pre = in.getdoubleParameter("P_pre");
    T_pre = in.getdoubleParameter("T_pre");
    M_wave = in.getdoubleParameter("M_wave");
}

// ------------------------------------------------------------------ Reaction
namespace {
int species_index(const Mechanism& me, const char* name) {
    for (int s = 0; s < me.NS; s++)
        if (me.species[s] == name) return s;
    throw Error(EU3D_ERR_ARG, std::string("ReactionInitial: the mechanism has no species ") + name);
}
}  // namespace

std::vector<double> Reaction::ReactionInitial(const Mesh& m, const Mechanism& me) {
    const int NS = me.NS, iO2 = species_index(me, "O2"), iH2 = species_index(me, "H2"),
              iN2 = species_index(me, "N2");
    const size_t NX = m.ni + 2 * m.bc, NY = m.nj + 2 * m.bc, NZ = m.nk + 2 * m.bc;
    const double cx = 0.038, cy = 0.0, r_in = 0.022, r_out = 1.2 * r_in, alpha = 5;
    // mass fractions of a mixture given by moles, the total summed in species order
    auto by_moles = [&](const std::vector<double>& mr) {
        double tot = 0.0;
        for (int s = 0; s < NS; s++) tot += mr[s] * me.mw[s];
        std::vector<double> y(NS);
        for (int s = 0; s < NS; s++) y[s] = mr[s] * me.mw[s] / tot;
        return y;
    };
    std::vector<double> mr_air(NS, 0.0), mr_fuel(NS, 0.0);
    mr_air[iO2] = 1.0;
    mr_air[iN2] = 3.76;
    mr_fuel[iH2] = 1.0;
    const std::vector<double> y_air = by_moles(mr_air), y_fuel = by_moles(mr_fuel);
    const double air_mw = me.mw[iO2] + 3.76 * me.mw[iN2];
    std::vector<double> Y(NX * NY * NZ * NS, 0.0);
    for (size_t i = 0; i < NX; i++)
        for (size_t j = 0; j < NY; j++) {
            const double r2 = std::pow(m.xnode[i] - cx, 2) + std::pow(m.ynode[j] - cy, 2);
            std::vector<double> y(NS, 0.0);
            if (r2 > std::pow(r_out, 2)) {
                y = y_air;
            } else if (r2 < std::pow(r_in, 2)) {
                y = y_fuel;
            } else {
                const double r = std::sqrt(r2);
                y[iH2] = std::exp(-alpha * std::pow((r - r_in) / (r_out - r_in), 2));
                y[iO2] = (1 - y[iH2]) * me.mw[iO2] / air_mw;
                y[iN2] = (1 - y[iH2]) * 3.76 * me.mw[iN2] / air_mw;
            }
            for (size_t k = 0; k < NZ; k++)
                for (int s = 0; s < NS; s++) Y[((i * NY + j) * NZ + k) * NS + s] = y[s];
        }
    return Y;
}

void Reaction::Trapezoid(Flowfield&) { ctx.check(eu3d_trapezoid(ctx.get()), "Trapezoid"); }

void Reaction::TrapezoidPrediction(Flowfield&, int nchem_max) {
    ctx.check(eu3d_trapezoid_prediction(ctx.get(), nchem_max), "TrapezoidPrediction");
}

void Reaction::TrapezoidAdaptive(Flowfield&, int nchem_max) {
    ctx.check(eu3d_trapezoid_adaptive(ctx.get(), nchem_max), "Trapezoid (adaptive)");
}

void Reaction::Diagonalized(Flowfield&) { ctx.check(eu3d_diagonalized(ctx.get()), "Diagonalized"); }

std::vector<double> Reaction::Nchem() const {
    std::vector<double> h(ctx.mesh().cells_with_ghosts());
    ctx.check(eu3d_get_field(ctx.get(), EU3D_FIELD_NCHEM, h.data()), "get Nchem");
    return h;
}

eu3d_chem_stats Reaction::stats() const {
    eu3d_chem_stats s{};
    ctx.check(eu3d_get_chem_stats(ctx.get(), &s), "chem stats");
    return s;
}

// ------------------------------------------------------------------ Control / Euler
void Control::FileRead(const std::string& ctrlfile) {
    FileReader c;
    if (!c.readFile(ctrlfile)) throw Error(EU3D_ERR_ARG, "cannot read control deck " + ctrlfile);
    Final_Time = c.getdoubleParameter("Final-Time");
    cfl = c.getdoubleParameter("cfl");
    if (cfl == 0) dt = c.getdoubleParameter("dt");
    count = c.getIntParameter("count");
    gridname = c.getStringParameter("gridname");
    initialization = c.findparameters("initialization") ? c.getStringParameter("initialization") : "";
    reaction_model = c.getStringParameter("reaction-model");
    Thermofile = c.getStringParameter("Thermofile");
    react = c.getStringParameter("reaction");
    if (react == "Trapezoid") Reaction_Sch = 0, type = 0;
    else if (react == "Trapezoid_Adaptive") Reaction_Sch = 0, type = 1;
    else if (react == "IMEX") Reaction_Sch = 1;
    else if (react == "DNN") Reaction_Sch = 2, type = 1;
    else throw Error(EU3D_ERR_ARG, "Invalid reaction scheme " + react);
    const std::string diff = c.getStringParameter("difference");
    if (diff == "MUSCL_1") Diff_Sch = 0;
    else if (diff == "MUSCL_2") Diff_Sch = 1;
    else throw Error(EU3D_ERR_UNSUPPORTED, "difference scheme " + diff);
    if (c.findparameters("m_block_x")) m_block_x = std::max(1, c.getIntParameter("m_block_x"));
    if (c.findparameters("m_block_y")) m_block_y = std::max(1, c.getIntParameter("m_block_y"));
    if (c.findparameters("m_block_z")) m_block_z = std::max(1, c.getIntParameter("m_block_z"));
    const std::string ta = c.getStringParameter("time-adv");
    if (ta == "EE") TimeAdv_Sch = TimeAdv::EE;
    else if (ta == "TVD_RK3") TimeAdv_Sch = TimeAdv::TVD_RK3;
    else throw Error(EU3D_ERR_ARG, "Invalid time-advance scheme " + ta);
}

Euler::Setup Euler::read_decks(const std::string& ctrlfile) {
    Setup s;
    s.c.FileRead(ctrlfile);
    s.m.MeshProcess(s.c.gridname);
    s.me.ReactionRead(s.c.reaction_model, s.c.Thermofile);
    return s;
}

Euler::Euler(const std::string& ctrlfile, uint32_t quirks) : Euler(read_decks(ctrlfile), quirks) {}

Euler::Euler(const Setup& s, uint32_t quirks) : Euler(s.c, s.m, s.me, quirks) {}

Euler::Euler(const Control& c, const Mesh& m, const Mechanism& me, uint32_t quirks)
    : mesh(m), mech(me), ctrl(c), ctx(m, me, quirks), TwoDim(ctx), React(ctx) {}

void Euler::set_surrogate(eu3d_mlp* m) { ctx.check(eu3d_set_surrogate(ctx.get(), m), "set_surrogate"); }

double Euler::now() {
    if (record) ctx.check(eu3d_sync(ctx.get()), "sync");
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void Euler::Trapezoid(int type) {
    const double t0 = record ? now() : 0;
    TwoDim.Update_after_Adv();
    const double t1 = record ? now() : 0;
    if (type == 0) {
        React.Trapezoid(TwoDim);
        TwoDim.Explicit();
    } else {
        React.TrapezoidAdaptive(TwoDim, nchem_max);
    }
    if (record) {
        phase.tr1 += t1 - t0;
        phase.tr2 += now() - t1;   // prediction and sub-steps run as one adaptive pass
    }
}

void Euler::IMEX() {
    const double t0 = record ? now() : 0;
    React.Diagonalized(TwoDim);
    const double t1 = record ? now() : 0;
    TwoDim.Update_IMEX(React);
    if (record) {
        phase.tr1 += t1 - t0;
        phase.tr2 += now() - t1;
    }
}

void Euler::Initialize() {
    if (ctrl.initialization.empty()) throw Error(EU3D_ERR_ARG, "the control deck names no initialization deck");
    InitialDeck in;
    in.InputRead(ctrl.initialization);
    TwoDim.FieldInitial(in, mesh, Reaction::ReactionInitial(mesh, mech));
}

void Euler::DNN(int type) {
    TwoDim.Update_after_Adv();
    if (type == 1) ctx.check(eu3d_surrogate_chem(ctx.get(), nchem_max, nullptr), "DNN (surrogate chemistry)");
}

void Euler::Computing(int count, const int* bc_kinds, int max_steps) {
    double time = 0.0;
    int num = 0;
    iteration = 0;
    const double Final_Time = ctrl.Final_Time;
    while (time < Final_Time && (max_steps <= 0 || iteration < max_steps)) {
        if (ctrl.cfl == 0) TwoDim.set_dt(ctrl.dt);
        else TwoDim.CFLcondition(ctrl.cfl, Final_Time);
        const double dt = TwoDim.dt();
        const double t1 = record ? now() : 0;
        if (time <= num * Final_Time / count && time + dt > num * Final_Time / count) {
            Output(num);
            num++;
        }
        const double t2 = record ? now() : 0;
        if (bc_kinds) TwoDim.FieldBoundary(bc_kinds);
        else TwoDim.FieldBoundary_3dRSBI();
        const double t3 = record ? now() : 0;
        TwoDim.Mpi_Boundary();
        const double t4 = record ? now() : 0;
        TwoDim.Advection(ctrl.TimeAdv_Sch, ctrl.Diff_Sch);
        const double t5 = record ? now() : 0;
        switch (ctrl.Reaction_Sch) {
        case 0: Trapezoid(ctrl.type); break;
        case 1: IMEX(); break;
        case 2: DNN(ctrl.type); break;
        }
        if (record) {
            phase.output += t2 - t1;
            phase.boundary += t3 - t2;
            phase.mpi_boundary += t4 - t3;
            phase.advection += t5 - t4;
            phase.reaction += now() - t5;
        }
        time += dt;
        iteration++;
        if (TwoDim.IsNan()) {
            Output(999);
            throw Error(EU3D_ERR_NONFINITE, "D or T is nan after iteration " + std::to_string(iteration));
        }
    }
}

void Euler::ComputingFused(int nsteps, const int bc_kinds[6], int rk, int nchem_max_) {
    eu3d_step_params p{};
    p.cfl = ctrl.cfl;
    p.final_time = ctrl.Final_Time > 0 ? ctrl.Final_Time : 1e30;
    for (int f = 0; f < 6; f++) p.bc_kinds[f] = bc_kinds[f];
    p.diff = ctrl.Diff_Sch;
    p.limiter = 0;
    p.rk = rk;
    switch (ctrl.Reaction_Sch) {
    case 0: p.chem = ctrl.type == 0 ? EU3D_CHEM_TRAPEZOID : EU3D_CHEM_TRAP_ADAPTIVE; break;
    case 1: p.chem = EU3D_CHEM_IMEX; break;
    default: p.chem = EU3D_CHEM_SURROGATE;
    }
    p.nchem_max = nchem_max_;
    if (ctrl.cfl == 0) TwoDim.set_dt(ctrl.dt);
    ctx.check(eu3d_step(ctx.get(), &p, nsteps), "eu3d_step");
    iteration += nsteps;
    if (TwoDim.IsNan()) throw Error(EU3D_ERR_NONFINITE, "D or T is nan");
}

// src/Euler.cpp:984-1021: the phase times of the run, and the loop time for comparing runs.
void Euler::Output_Total_Time(double file_reader, double mesh_t, double input, double reaction_read,
                              double field_initial, double computing, double output) const {
    const std::string base = outdir + "/output_1/time_record/";
    make_dirs(base + "total");
    make_dirs(base + "compare");
    std::ofstream t(base + "total/Total_Time_0.txt");
    t << "myid:" << '\t' << 0 << '\n';
    t << "FileReader:" << '\t' << file_reader << '\n';
    t << "Mpi_Initial:" << '\t' << 0.0 << '\n';
    t << "MeshProcess:" << '\t' << mesh_t << '\n';
    t << "InputRead:" << '\t' << input << '\n';
    t << "ReactionRead:" << '\t' << reaction_read << '\n';
    t << "FieldInitial:" << '\t' << field_initial << '\n';
    t << "MPI_Barrier:" << '\t' << 0.0 << '\n';
    t << "Computing:" << '\t' << computing << '\n';
    t << "Output:" << '\t' << output << '\n';
    if (!t) throw Error(EU3D_ERR_ARG, "cannot write " + base + "total/Total_Time_0.txt");
    std::ofstream c(base + "compare/" + ctrl.react + "_mpi_1_omp_1.txt");
    c << computing << '\t' << 0.0;
    std::cout << "Loop time: " << computing << std::endl;
}

// src/Euler.cpp:1022-1099: the loop's phases, per reaction scheme, and the per-rank files.
void Euler::Output_Compute_Time() const {
    const std::string base = outdir + "/output_1/time_record/";
    make_dirs(base + "compute");
    make_dirs(base + "all");
    std::ofstream o(base + "compute/Compute_Time_0.txt");
    o << "myid:" << '\t' << 0 << '\n';
    o << "FieldBoundary:" << '\t' << phase.boundary << '\n';
    o << "Mpi_Boundary:" << '\t' << phase.mpi_boundary << '\n';
    o << "Advection:" << '\t' << phase.advection << '\n';
    o << "Reaction:" << '\t' << phase.reaction << '\n';
    o << "Output:" << '\t' << phase.output << '\n';
    switch (ctrl.Reaction_Sch) {
    case 0:
        o << "Update_after_Advection:" << '\t' << phase.tr1 << '\n';
        o << "Trapezoid_Prediction:" << '\t' << phase.prediction << '\n';
        o << "Trapezoid:" << '\t' << phase.tr2 << '\n';
        o << "DLB:" << '\t' << 0.0 << '\n';
        o << "Wait:" << '\t' << 0.0 << '\n';
        break;
    case 1:
        o << "Diagnalized:" << '\t' << phase.tr1 << '\n';
        o << "Update_IMEX:" << '\t' << phase.tr2 << '\n';
        o << "FileBoundary:" << phase.boundary << " " << 0.0 << " " << 0.0 << " " << 0.0 << " " << 0.0 << " " << 0.0
          << " " << 0.0 << '\n';
        break;
    }
    if (!o) throw Error(EU3D_ERR_ARG, "cannot write " + base + "compute/Compute_Time_0.txt");
    std::ofstream(base + "all/All_wait.txt") << 0.0 << '\n';
    std::ofstream(base + "all/All_react.txt") << phase.tr1 + phase.tr2 + phase.prediction << '\n';
    std::ofstream(base + "all/All_advert.txt") << phase.advection << '\n';
    std::ofstream(base + "all/All_DLB.txt") << 0.0 << '\t' << 0.0 << '\t' << 0.0 << '\t' << 0.0 << '\t' << 0.0 << '\t'
                                            << 0.0 << '\t' << 0.0 << '\n';
}

void Euler::Output(int num) {
    const std::vector<double> nch = React.Nchem();
    const std::vector<double> T = TwoDim.field(EU3D_FIELD_T), D = TwoDim.field(EU3D_FIELD_D),
                              P = TwoDim.field(EU3D_FIELD_P), U = TwoDim.field(EU3D_FIELD_U),
                              V = TwoDim.field(EU3D_FIELD_V), Ma = TwoDim.field(EU3D_FIELD_MA),
                              Yi = TwoDim.field(EU3D_FIELD_YI);
    const int bx = ctrl.m_block_x, by = ctrl.m_block_y, bz = ctrl.m_block_z, nb = bx * by * bz;
    if (nb <= 1) {
        WriteTecplot(outdir + "/Series" + ctrl.react + std::to_string(num) + "_0.dat", mesh, T, D, P, U, V, Ma, Yi,
                     mech.NS, nch, 0.0);
        return;
    }
    const std::string dir = outdir + "/output_" + std::to_string(nb);
    make_dirs(dir);
    OutputBlock b;
    b.ni = (mesh.ni - 1) / bx + 1;
    b.nj = (mesh.nj - 1) / by + 1;
    b.nk = (mesh.nk - 1) / bz + 1;
    for (int id = 0; id < nb; id++) {
        b.i0 = (id % bx) * (b.ni - 1);
        b.j0 = (id % (bx * by)) / bx * (b.nj - 1);
        b.k0 = id / (bx * by) * (b.nk - 1);
        WriteTecplot(dir + "/results" + ctrl.react + std::to_string(num) + "_" + std::to_string(id) + ".dat", mesh, T,
                     D, P, U, V, Ma, Yi, mech.NS, nch, 0.0, &b);
    }
}

void WriteTecplot(const std::string& filename, const Mesh& m, const std::vector<double>& T, const std::vector<double>& D,
                  const std::vector<double>& P, const std::vector<double>& U, const std::vector<double>& V,
                  const std::vector<double>& Ma, const std::vector<double>& Yi, int ns, const std::vector<double>& Nchem,
                  double wait_time, const OutputBlock* blk) {
    OutputBlock b;
    if (blk) b = *blk;
    else b.ni = m.ni, b.nj = m.nj, b.nk = m.nk;
    if (b.i0 < 0 || b.j0 < 0 || b.k0 < 0 || b.ni < 1 || b.nj < 1 || b.nk < 1 || b.i0 + b.ni > m.ni ||
        b.j0 + b.nj > m.nj || b.k0 + b.nk > m.nk)
        throw Error(EU3D_ERR_ARG, "WriteTecplot: block outside the grid");
    std::ofstream out(filename);
    if (!out.is_open()) throw Error(EU3D_ERR_ARG, "cannot write " + filename);
    const int NY = m.nj + 2 * m.bc, NZ = m.nk + 2 * m.bc;
    out << "Variables = X,Y,Z,T,D,P,U,V,Ma,Y1,Y2,Y3,Y4,Y5,Y6,Y7,L,WaitTime\n";
    out << "ZONE I=" << b.ni << '\t' << "J=" << b.nj << '\t' << "K=" << b.nk << '\n';
    out << "datapacking=block\n";
    // one block per variable, i fastest over the block's nodes; reference layout index ((i*NY)+j)*NZ+k
    const int ib = m.bc + b.i0, jb = m.bc + b.j0, kb = m.bc + b.k0;
    auto block = [&](auto value) {
        for (int k = kb; k < kb + b.nk; k++)
            for (int j = jb; j < jb + b.nj; j++)
                for (int i = ib; i < ib + b.ni; i++) out << value(i, j, k) << '\n';
    };
    auto cell = [&](int i, int j, int k) { return ((size_t)i * NY + j) * NZ + k; };
    block([&](int i, int, int) { return m.xnode[i]; });
    block([&](int, int j, int) { return m.ynode[j]; });
    block([&](int, int, int k) { return m.znode[k]; });
    for (const std::vector<double>* f : {&T, &D, &P, &U, &V, &Ma})
        block([&](int i, int j, int k) { return (*f)[cell(i, j, k)]; });
    for (int s = 0; s < 7; s++) block([&](int i, int j, int k) { return Yi[cell(i, j, k) * ns + s]; });
    block([&](int i, int j, int k) { return Nchem[cell(i, j, k)]; });
    block([&](int, int, int) { return wait_time; });
}

}  // namespace eu3d_host
