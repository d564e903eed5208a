// Compiles the parts of the reference CPU path that build from their own
// sources (no CUDA headers needed) and exposes them through a C ABI so the
// oracle's restatements can be checked against them bit for bit.
// Built by oracle/build_ref.sh into oracle/_ref/libref.so; test-only.
//
//   MUSCL_2 / Minmod      include/Differ.hpp:33-43, 149-238
//   Function::sum         src/Function.cpp:17-42
//   FileReader            src/FileReader.cpp:53-84, include/FileReader.hpp:74-111
//   Mesh::MeshProcess     src/Mesh.cpp:36-142 (single block)
#include <cstring>
#include <string>
#include "Differ.hpp"
#include "Function.hpp"
#include "FileReader.hpp"
#include "Mesh.hpp"

// Mesh keeps its node arrays private and befriends class Euler
// (include/Mesh.hpp:57); this accessor takes that name to read them.
class Euler
{
  public:
    static void nodes(Mesh &m, Array<double, 1> &x, Array<double, 1> &y, Array<double, 1> &z)
    {
        x = m.xnode;
        y = m.ynode;
        z = m.znode;
    }
};

extern "C" {

void ref_muscl2(int direction, int NI, int NJ, int NK, int bc, const double *fi, double *FLR)
{
    Array<double, 1> x(NI), y(NJ), z(NK);
    Array<double, 3> f(NI, NJ, NK);
    Array<double, 4> flr(NI, NJ, NK, 2);
    std::memcpy(f.Getbuf(), fi, sizeof(double) * (size_t)NI * NJ * NK);
    std::memcpy(flr.Getbuf(), FLR, sizeof(double) * (size_t)NI * NJ * NK * 2);
    MUSCL_2(direction, x, y, z, flr, f, bc, 0);
    std::memcpy(FLR, flr.Getbuf(), sizeof(double) * (size_t)NI * NJ * NK * 2);
}

double ref_minmod(double r) { return Minmod(r); }

double ref_sum(int flag, int n, const double *a, const double *b)
{
    Array<double, 1> A(n), B(n);
    std::memcpy(A.Getbuf(), a, sizeof(double) * n);
    std::memcpy(B.Getbuf(), b, sizeof(double) * n);
    Function F;
    return F.sum(flag, A, B);
}

// kind 0: int, 1: double, 2: string
int ref_filereader(const char *path, const char *key, int kind, int *ival, double *dval, char *sval, int slen)
{
    FileReader R;
    R.readFile(path);
    if (!R.findparameters(key))
        return -1;
    if (kind == 0)
        *ival = R.getIntParameter(key);
    else if (kind == 1)
        *dval = R.getdoubleParameter(key);
    else
        std::snprintf(sval, slen, "%s", R.getStringParameter(key).c_str());
    return 0;
}

int ref_mesh(const char *gridfile, int bc, double *xnode, double *ynode, double *znode, int *sizes)
{
    m_block_x = m_block_y = m_block_z = 1;
    myid = myid_x = myid_y = myid_z = 0;
    numprocs = 1;
    Mesh M;
    char name[1024];
    std::snprintf(name, sizeof name, "%s", gridfile);
    M.MeshProcess(name);
    Array<double, 1> X(1), Y(1), Z(1);
    Euler::nodes(M, X, Y, Z);
    sizes[0] = (int)X.GetSize();
    sizes[1] = (int)Y.GetSize();
    sizes[2] = (int)Z.GetSize();
    (void)bc;
    if (xnode) std::memcpy(xnode, X.Getbuf(), sizeof(double) * X.GetSize());
    if (ynode) std::memcpy(ynode, Y.Getbuf(), sizeof(double) * Y.GetSize());
    if (znode) std::memcpy(znode, Z.Getbuf(), sizeof(double) * Z.GetSize());
    return 0;
}
}
