// eu3d_host.hpp -- host side of the MI355X reactive-Euler step: the
// reference's Flowfield / Reaction / TimeAdv / Mesh / Euler class interfaces
// (include/Flowfield.hpp, include/Reaction.hpp, include/TimeAdv.hpp,
// include/Mesh.hpp, include/Euler.hpp) and its input formats, driving the
// device through the C ABI of include/eu3d.h.
//
// Host arrays keep the reference layout (include/Array.hpp:33-43): row-major
// over the padded grid, last index fastest.  The device holds the solve; the
// host copies are filled on demand (Euler::SyncOutput) for dumps.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/eu3d.h"

namespace eu3dh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// FileReader (src/FileReader.cpp:53-84, include/FileReader.hpp:74-111): the
// first two tokens of each line, kept across lines that run short; numbers
// re-stored at 6 significant digits; '#' skips a line only at token start.
class FileReader {
  public:
    bool readFile(const std::string &name);
    bool has(const std::string &key) const { return parameters.count(key) != 0; }
    int getIntParameter(const std::string &key) const;
    double getdoubleParameter(const std::string &key) const;
    std::string getStringParameter(const std::string &key) const;
    double getdoubleDefault(const std::string &key, double dflt) const
    {
        return has(key) ? getdoubleParameter(key) : dflt;
    }

  private:
    std::map<std::string, std::string> parameters;
};

// Mesh (src/Mesh.cpp:36-142), one block.
class Mesh {
  public:
    void MeshProcess(const std::string &gridname);
    int total_ni = 0, total_nj = 0, total_nk = 0;
    int ni = 0, nj = 0, nk = 0, bc = 0;
    double xa = 0, xb = 0, ya = 0, yb = 0, za = 0, zb = 0, dx = 0, dy = 0, dz = 0;
    std::vector<double> xnode, ynode, znode;
};

// Reaction (src/Reaction.cpp): mechanism/thermo readers, NASA-9 thermo,
// RSBI composition.
class Reaction {
  public:
    void ReactionRead(const std::string &reaction_model, const std::string &thermofile);
    double GetCpi(double T, double R, int SP) const;
    double GetHi(double T, double R, int SP) const;
    double GetSi(double T, double R, int SP) const;
    eu3d_mechanism Mechanism() const;

    int NS = 0, NR = 0, TB = 0;
    const double R = 8.31434;  // J/(mol K)
    const double Ru = 1.987;   // cal/(mol K)
    const double P0 = 101325;  // Pa
    std::vector<double> Stoi_F, Stoi_B, React_TB; // [NS*NR], s*NR + r
    std::vector<double> Af, Bf, Eaf;              // [NR]
    std::vector<double> Coeff0, Coeff1;           // [9*NS], k*NS + s
    std::vector<double> Mw, Ri;                   // [NS]
    std::vector<double> Mr_temp;                  // [NS]
};

// Flowfield (src/Flowfield.cpp): the field arrays, the initial conditions and
// the step operations, which run on the device.
class Flowfield {
  public:
    void InputRead(const std::string &initialization, const Mesh &mesh);
    void Construction();
    void FieldInitial(const Reaction &react);
    void Upload();
    void FieldBoundary_3dRSBI();
    void CFLcondition(double cfl, double Final_Time);
    void Advection(int time_adv);
    void Update_after_Adv();
    void Download();

    // geometry
    int ni = 0, nj = 0, nk = 0, bc = 0, NI = 0, NJ = 0, NK = 0;
    size_t N = 0;
    const Mesh *mesh = nullptr;
    int NS = 0, NR = 0;
    double dt = 0, time = 0;
    // input keys
    double P_pre = 0, T_pre = 0, M_wave = 0;
    std::string case_name;
    FileReader input;
    // fields
    std::vector<double> U, V, W, P, D, T, C, Ma, Wav, Rgas, Cp, H, E, Gamma;
    std::vector<double> Mc, Mi, Yi, Di, CMS, CS;
    std::vector<int> Nchem;
    eu3d_ctx *ctx = nullptr;
    // ghost-fill set (EU3D_BOUNDARY_*) and, for ODW, the inflow state
    // P rho u v w h gamma Y[NS] T E taken from the initial free stream
    int boundary = EU3D_BOUNDARY_RSBI;
    std::vector<double> inflow;

    size_t idx(int i, int j, int k) const { return ((size_t)i * NJ + j) * NK + k; }
    eu3d_state HostState();

  private:
    void ic_rsbi(const Reaction &react);
    void ic_sbi(const Reaction &react);
    void ic_znd(const Reaction &react);
    void ic_odw(const Reaction &react);
};

// TimeAdv (src/TimeAdv.cpp): EE and TVD_RK3 on the device.
class TimeAdv {
  public:
    static void EE(eu3d_ctx *ctx, double dt);
    static void TVD_RK3(eu3d_ctx *ctx, double dt);
};

// Euler (src/Euler.cpp): ctrl file, the time loop, output.
class Euler {
  public:
    // device=false reads the inputs and builds the initial state on the host
    // only (no GPU needed); Attach() then creates the device context.
    explicit Euler(const std::string &ctrlfile, const std::string &base_dir = "", bool device = true);
    void Attach();
    ~Euler();
    void Computing(int max_steps);
    void SyncOutput();
    void WriteStateDump(const std::string &path, int num);
    void ReadStateDump(const std::string &path, double restart_time);
    // Euler::Output (src/Euler.cpp:1536-1795): Tecplot Series file of output
    // number num under <out_dir>/output_1/results, the state dump when
    // EU3D_DUMP_STATE=1 and num == count, and the load logs so far.
    void Output(int num);
    void SetOutput(const std::string &dir) { out_dir = dir; }
    void WriteLoadLogs();

    std::string out_dir;    // empty: no scheduled output
    std::string react_name; // ctrl "reaction" keyword (file names)
    int out_num = 0;        // next scheduled output (src/Euler.cpp:313)
    std::vector<int> NchemMax;        // Load_Loop.dat (Reaction::GetNchemMax)
    std::vector<int> LoadIterations;  // LoadDegree.dat rows
    void SetChemFlags(int f) { chem_flags = f; }

    std::string base;
    double Final_Time = 0, cfl = 0, dt = 0;
    int count = 0, TimeAdv_Sch = 0, iteration = 0;
    int Reaction_type = 1; // ctrl "reaction": Trapezoid = 0, Trapezoid-Adaptive = 1, IMEX = 2, DNN/Cantera = 3
                           // (no chemistry) (src/Euler.cpp:136-160)
    int Diff_Sch = 2;      // ctrl "difference": MUSCL_1 = 1, MUSCL_2 = 2 (src/Euler.cpp:162-176)
    int chem_flags = EU3D_CHEM_COMPACT | EU3D_CHEM_KEEP_CMS;
    double time = 0;
    Mesh Mymesh;
    Reaction React;
    Flowfield TwoDim;

  private:
    std::string path(const std::string &p) const;
};

void check(int rc, eu3d_ctx *ctx, const char *what);

} // namespace eu3dh
