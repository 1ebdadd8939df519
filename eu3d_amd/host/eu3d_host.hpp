// C++ host layer over the eu3d C ABI, keeping the reference's class and method names so that a
// driver written against the reference (Flowfield / Reaction / TimeAdv / Euler, FileReader,
// Mesh) maps one-to-one. Each method body is one or a few ABI calls on the context that owns the
// device state; the arrays the reference driver reads as a friend are downloaded on demand in
// the reference's Array layout (row-major, last index fastest, ghosts included).
//
// Reference interfaces mirrored (paths relative to the reference tree):
//   FileReader   include/FileReader.hpp:15-56, src/FileReader.cpp:53-84
//   Mesh         src/Mesh.cpp:75-121 (one block: m_block_x = m_block_y = m_block_z = 1)
//   Reaction     include/Reaction.hpp:143-193, ReactionRead src/Reaction.cpp:150-326
//   Flowfield    include/Flowfield.hpp:142-211
//   TimeAdv      include/TimeAdv.hpp:20-23
//   Euler        include/Euler.hpp:112-140, src/Euler.cpp:30-333,748-760,792-973
//
// Errors: the reference prints and calls abort(); here every failure throws eu3d_host::Error
// carrying the ABI status, and Euler::Computing reproduces the reference's NaN abort path by
// writing output 999 and throwing.
#pragma once

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eu3d.h"

namespace eu3d_host {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code(code) {}
    int code;
};

// Key/value control decks (ctrl, grid, input files). Parse rules of the reference reader: a
// line's first two whitespace tokens are key and value (tokens carry over from the previous line
// when a line has fewer); lines whose key or value starts with '#' are skipped; a value that
// begins with a number is stored as that number printed with 6 significant digits, otherwise as
// text. getIntParameter / getdoubleParameter return 0 when the stored text does not start with
// an integer (the reference's stoi pre-check).
class FileReader {
  public:
    bool readFile(const std::string& name);
    int getIntParameter(const std::string& key) const;
    double getdoubleParameter(const std::string& key) const;
    std::string getStringParameter(const std::string& key) const;
    bool findparameters(const std::string& key) const;
    void setParameter(const std::string& key, const std::string& value);
    void setParameter(const std::string& key, double value);
    const std::map<std::string, std::string>& all() const { return parameters; }

  private:
    const std::string& at(const std::string& key) const;
    std::map<std::string, std::string> parameters;
};

// Vertex grid of one block with bc ghost nodes per side (Mesh::Initial + Mesh::Boundary).
struct Mesh {
    int ni = 0, nj = 0, nk = 0, bc = 2;
    double dx = 0, dy = 0, dz = 0;
    std::vector<double> xnode, ynode, znode;
    void MeshProcess(const std::string& gridfile);           // total_n*, a/b extents, bc from a deck
    void Initial(int total_ni, double xa, double xb, int total_nj, double ya, double yb, int total_nk,
                 double za, double zb, int bc);
    void Uniform(int ni, int nj, int nk, double d, int bc = 2);   // nodes 0, d, 2d, ... per axis
    size_t cells_with_ghosts() const { return (size_t)(ni + 2 * bc) * (nj + 2 * bc) * (nk + 2 * bc); }
};

// Mechanism and thermo tables as Reaction::ReactionRead leaves them.
struct Mechanism {
    int NS = 0, NR = 0, TB = 0;
    std::vector<std::string> species;
    std::vector<double> stoi_f, stoi_b, react_tb;   // [NS * NR], (s, r) row-major
    std::vector<double> af, bf, eaf;                // [NR]
    std::vector<double> coeff0, coeff1;             // [9 * NS], (k, s) row-major
    std::vector<double> mw, mr;                     // [NS]
    void ReactionRead(const std::string& reaction_model, const std::string& thermofile);
    eu3d_mech view() const;
};

// One device context (the state of one Flowfield + Reaction pair).
class Context {
  public:
    Context(const Mesh& m, const Mechanism& mech, uint32_t quirks);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    eu3d_ctx* get() const { return ctx; }
    void check(int rc, const char* what) const;
    const Mesh& mesh() const { return m; }

  private:
    eu3d_ctx* ctx = nullptr;
    Mesh m;
};

// Time-advance scheme ids of the control deck ("time-adv": EE, TVD_RK3). EE is applied inside
// Flowfield::Advection as in the reference; TVD_RK3 is the reference's no-op (RHS = RHS,
// src/TimeAdv.cpp:57-67) and is refused rather than silently reproduced.
class TimeAdv {
  public:
    enum Scheme { EE = 0, TVD_RK3 = 1 };
};

// Shock deck values of Flowfield::InputRead (src/Flowfield.cpp:33-58): pre-shock pressure [Pa]
// and temperature [K] and the Mach number of the moving shock.
struct InitialDeck {
    double P_pre = 0, T_pre = 0, M_wave = 0;
    void InputRead(const std::string& initialization);
};

class Reaction;

class Flowfield {
  public:
    explicit Flowfield(Context& c) : ctx(c) {}
    // FieldInitial's second loop + GetPartial_T from the primitive fields T, P, U, V, W, Yi set
    // over every cell (SetPrimitives), src/Flowfield.cpp:193-229.
    void SetPrimitives(const std::vector<double>& T, const std::vector<double>& P, const std::vector<double>& U,
                       const std::vector<double>& V, const std::vector<double>& W, const std::vector<double>& Yi);
    void FieldInitial();
    // The whole of FieldInitial (src/Flowfield.cpp:154-229): the moving shock's state first
    // (at rest at P_pre, T_pre where x >= 0.01 m, the post-shock state of the Mach M_wave shock
    // behind it, V = W = 0, over every cell), then the second loop above from those and Yi.
    void FieldInitial(const InitialDeck& in, const Mesh& m, const std::vector<double>& Yi);
    void CFLcondition(double cfl, double Final_Time);
    void FieldBoundary_3dRSBI();
    void FieldBoundary(const int kinds[6]);
    void Mpi_Boundary() {}   // one GPU holds the whole domain: no halo exchange
    // Diff_Sch 0 = MUSCL_1, 1 = MUSCL_2; limit as include/Differ.hpp:188-220 (the reference
    // hard-wires 0, src/Flowfield.cpp:1336-1352). fused selects the two-pass kernels.
    void Advection(int TimeAdv_Sch, int Diff_Sch, int limit = 0, bool fused = true);
    void Update_after_Adv();
    void Explicit();
    void Update_IMEX(Reaction& React);
    void Update_after_CS();
    void GetPartial_T();
    void set_dt(double dt);
    double dt() const;
    double time() const;
    // Friend-read state of the reference driver, downloaded in reference layout.
    std::vector<double> field(int id) const;
    void set_field(int id, const std::vector<double>& host);
    bool IsNan() const;   // D or T non-finite

  private:
    friend class Reaction;
    Context& ctx;
};

class Reaction {
  public:
    explicit Reaction(Context& c) : ctx(c) {}
    // ReactionInitial (src/Reaction.cpp:340-397): the mass fractions of the 3dRSBI start over
    // every cell including ghosts, [cells_with_ghosts * NS] in reference layout. A cylinder
    // (axis along z through x = 0.038 m, y = 0) of radius 0.022 m holds pure H2; outside 1.2
    // times that radius is air (O2 + 3.76 N2 by moles); between them H2 falls off as
    // exp(-5 ((r - r_i) / (r_o - r_i))^2) and O2 / N2 share the rest in the air's mass ratio. The
    // reference indexes O2, H2 and N2 as species 4, 5 and 8 of its mechanism; here they are
    // found by name, which is the same for that mechanism.
    static std::vector<double> ReactionInitial(const Mesh& m, const Mechanism& me);
    // TwoDim.Di = React.Trapezoid(TwoDim.Mc, TwoDim.Di, TwoDim.T, TwoDim.dt): full field,
    // one step of dt (src/Reaction.cpp:412-485); the caller follows with Explicit().
    void Trapezoid(Flowfield& f);
    // TrapezoidPrediction + the per-cell Trapezoid / Explicit loop (src/Euler.cpp:526-556), with
    // active-cell compaction and the persistent kernel in place of the load balancer.
    void TrapezoidPrediction(Flowfield& f, int nchem_max = 0);
    void TrapezoidAdaptive(Flowfield& f, int nchem_max = 0);
    void Diagonalized(Flowfield& f);
    std::vector<double> Nchem() const;
    eu3d_chem_stats stats() const;

  private:
    Context& ctx;
};

// Control deck values Euler::FileRead keeps (src/Euler.cpp:73-190).
struct Control {
    double Final_Time = 0, cfl = 0, dt = 0;
    int count = 1;
    std::string gridname, initialization, reaction_model, Thermofile;
    int Reaction_Sch = 1;   // 0 Trapezoid, 1 IMEX, 2 DNN (Global.hpp scheme ids)
    int type = 0;           // Trapezoid: 0 full field, 1 adaptive; DNN: 1 surrogate
    int Diff_Sch = 1;       // 0 MUSCL_1, 1 MUSCL_2
    int TimeAdv_Sch = 0;    // EE
    // MPI blocks of the reference run (m_block_x/y/z, src/Euler.cpp:197-215). One GPU holds the
    // whole domain; Output writes one file per block as the reference's ranks do.
    int m_block_x = 1, m_block_y = 1, m_block_z = 1;
    std::string react = "IMEX";
    void FileRead(const std::string& ctrlfile);
};

class Euler {
  public:
    // Reads the control deck, the grid deck and the mechanism / thermo files it names, and
    // creates the device context (the reference constructor up to FieldInitial).
    Euler(const std::string& ctrlfile, uint32_t quirks = EU3D_QUIRKS_FIXED);
    Euler(const Control& c, const Mesh& m, const Mechanism& mech, uint32_t quirks = EU3D_QUIRKS_FIXED);
    // The reference loop (src/Euler.cpp:266-333): CFL dt, output check, ghost fill, advection,
    // chemistry, NaN check; each method one ABI call. bc_kinds replace FieldBoundary_3dRSBI
    // when given. Stops at Final_Time or after max_steps steps (0: no limit).
    void Computing(int count, const int* bc_kinds = nullptr, int max_steps = 0);
    // The same step composition run as the device loop of eu3d_step (no host round trip per step).
    void ComputingFused(int nsteps, const int bc_kinds[6], int rk = 1, int nchem_max = 0);
    void Trapezoid(int type);
    void IMEX();
    void DNN(int type);
    // The reference constructor's initial state (src/Euler.cpp:56-68): the shock deck named by
    // the control deck's "initialization", ReactionInitial, then FieldInitial.
    void Initialize();
    // Tecplot block file as src/Euler.cpp:792-973 writes it. One block: <out>/Series<react><num>_0.dat.
    // With m_block_x * m_block_y * m_block_z > 1, one file per block as the reference's ranks
    // write them, <out>/output_<blocks>/results<react><num>_<id>.dat, with
    // id = bx + m_block_x (by + m_block_y bz) (src/Euler.cpp:213-215) and each block's nodes
    // (Mesh::Initial, src/Mesh.cpp:75-96: neighbouring blocks share their end nodes).
    void Output(int num);
    void set_output_dir(const std::string& d) { outdir = d; }
    void set_surrogate(eu3d_mlp* m);
    // Time records (src/Euler.cpp:984-1099). With record on, Computing times each phase of the
    // loop (the device is synchronised around it, so record only when the numbers are wanted).
    // Output_Total_Time / Output_Compute_Time then write <outdir>/output_1/time_record/... in
    // the reference's formats, as its rank 0 of one rank: MPI phases, load balancing and waits
    // are 0, and the boundary-fill split of the IMEX record holds the whole fill in its first
    // entry.
    void set_time_record(bool on) { record = on; }
    struct PhaseTimes {
        double boundary = 0, mpi_boundary = 0, advection = 0, reaction = 0, output = 0;
        double tr1 = 0, tr2 = 0, prediction = 0;   // Update_after_Adv / Diagonalized, the rest
    };
    PhaseTimes phase;
    void Output_Total_Time(double file_reader, double mesh, double input, double reaction_read,
                           double field_initial, double computing, double output) const;
    void Output_Compute_Time() const;

    Mesh mesh;
    Mechanism mech;
    Control ctrl;
    Context ctx;
    Flowfield TwoDim;
    Reaction React;
    int iteration = 0;

  private:
    struct Setup {
        Control c;
        Mesh m;
        Mechanism me;
    };
    static Setup read_decks(const std::string& ctrlfile);
    Euler(const Setup& s, uint32_t quirks);
    std::string outdir = ".";
    int nchem_max = 0;
    bool record = false;
    double now();   // seconds, after the device has finished queued work when recording
};

// Tecplot block writer: variables X, Y, Z, T, D, P, U, V, Ma, Y1..Y7, L (Nchem), WaitTime over
// the interior, each a block of values in i-fastest order, printed as std::ostream prints a
// double (6 significant digits).
// With blk, only the interior nodes [i0, i0 + ni) x [j0, j0 + nj) x [k0, k0 + nk) (interior
// indices, from 0) are written, as the block that holds them writes its file.
struct OutputBlock {
    int i0 = 0, j0 = 0, k0 = 0, ni = 0, nj = 0, nk = 0;
};
void WriteTecplot(const std::string& filename, const Mesh& m, const std::vector<double>& T, const std::vector<double>& D,
                  const std::vector<double>& P, const std::vector<double>& U, const std::vector<double>& V,
                  const std::vector<double>& Ma, const std::vector<double>& Yi, int ns, const std::vector<double>& Nchem,
                  double wait_time, const OutputBlock* blk = nullptr);

}  // namespace eu3d_host

// The deck-driven driver of the eu3d_run program, in-process: argv as the program takes it
// (argv[0] is ignored). Writes its report to stdout; returns the program's exit status.
extern "C" int eu3d_run_main(int argc, char** argv);
