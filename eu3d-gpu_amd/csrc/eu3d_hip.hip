// eu3d_hip.hip -- gfx950 kernels and the C ABI (include/eu3d.h) of the
// reactive-Euler step: ghost fill, fused MUSCL+AUSM+conservative update,
// Newton state recovery, CFL reduction and adaptive trapezoid chemistry with
// stiff-cell compaction.
//
// Device layout (structure of arrays over the padded grid, k fastest):
//   prim[f*N + idx], f = P D U V W H Gamma Y0..Y(ns-1)
//   T, E, C [N];  CS[n*N + idx], n < ns+4;  CS0 (RK3 stage 0);  CMS[s*N + idx]
//   nchem[N] (int)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eu3d.h"
#include "eu3d_device.h"

using namespace eu3d;

namespace {

constexpr int kBlock = 256;

struct ChemCounters {
    unsigned int queue_count;
    unsigned int fetch;
    unsigned long long invalid;
    unsigned long long nonfinite;
    unsigned long long stiff;
    unsigned long long substeps;
    int max_nchem;
    int pad;
    unsigned long long cfl_bits;
};

struct TimedKind {
    hipEvent_t a, b;
    int kind;
};

} // namespace

struct eu3d_ctx {
    Grid g;
    int ns, nr, nv, nprim, ncu;
    long long n_int;
    hipStream_t stream;
    DevMech hmech;
    DevMech *mech;
    double *x, *y, *z;
    double *prim, *T, *E, *C, *CS, *CS0, *CMS, *scratch;
    int *nchem, *queue;
    ChemCounters *cnt;
    bool has_state, cms_valid;
    bool generated; // tables bit-identical to MechH2JA919: use the compile-time kernels
    bool generated_available;
    std::string err;
    bool timing;
    std::vector<TimedKind> pending;
    std::vector<hipEvent_t> free_events;
    double last_ms[5], total_ms[5];
    long long calls[5];
};

#define CK(call)                                                                                  \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
            return EU3D_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

#define FAIL(code, msg)                                                                           \
    do {                                                                                          \
        ctx->err = (msg);                                                                         \
        return (code);                                                                            \
    } while (0)

enum { K_FLUX = 0, K_RECOVER = 1, K_CHEM = 2, K_GHOST = 3, K_CFL = 4 };

// ---------------------------------------------------------------------------
// timing (HIP events on the context stream)
// ---------------------------------------------------------------------------
static hipEvent_t get_event(eu3d_ctx *ctx)
{
    if (!ctx->free_events.empty()) {
        hipEvent_t e = ctx->free_events.back();
        ctx->free_events.pop_back();
        return e;
    }
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

struct Timer {
    eu3d_ctx *ctx;
    int kind;
    hipEvent_t a = nullptr;
    Timer(eu3d_ctx *c, int k) : ctx(c), kind(k)
    {
        if (ctx->timing) {
            a = get_event(ctx);
            if (a)
                (void)hipEventRecord(a, ctx->stream);
        }
    }
    ~Timer()
    {
        if (ctx->timing && a) {
            hipEvent_t b = get_event(ctx);
            if (b) {
                (void)hipEventRecord(b, ctx->stream);
                ctx->pending.push_back({a, b, kind});
            }
        }
    }
};

static void resolve_timing(eu3d_ctx *ctx)
{
    for (auto &t : ctx->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(t.b) == hipSuccess && hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            ctx->last_ms[t.kind] = ms;
            ctx->total_ms[t.kind] += ms;
            ctx->calls[t.kind] += 1;
        }
        ctx->free_events.push_back(t.a);
        ctx->free_events.push_back(t.b);
    }
    ctx->pending.clear();
}

// ---------------------------------------------------------------------------
// index helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool interior_cell(const Grid &g, long long t, int &i, int &j, int &k)
{
    long long nk = g.nk, nj = g.nj;
    if (t >= (long long)g.ni * nj * nk)
        return false;
    k = (int)(t % nk) + g.bc;
    long long r = t / nk;
    j = (int)(r % nj) + g.bc;
    i = (int)(r / nj) + g.bc;
    return true;
}

static unsigned int blocks_for(long long n, int b = kBlock) { return (unsigned int)((n + b - 1) / b); }

// The mechanism seen by a kernel: the run-time tables, or a generated type
// whose reaction tables are static constexpr members (it keeps the run-time
// pointer for the thermo tables).
template <class MT>
__device__ __forceinline__ decltype(auto) mech_ref(const DevMech *p)
{
    if constexpr (MT::kGenerated)
        return MT{p};
    else
        return static_cast<const DevMech &>(*p);
}

// ---------------------------------------------------------------------------
// ghost fill (src/Flowfield.cpp:593-807)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void copy_point(const Grid &g, int nprim, double *prim, double *T, double *E,
                                           long long d, long long s, int flip_field)
{
    for (int f = 0; f < nprim; f++) {
        double v = prim[f * g.N + s];
        prim[f * g.N + d] = (f == flip_field) ? -v : v;
    }
    T[d] = T[s];
    E[d] = E[s];
}

// x faces: left mirror copy (freestream), right mirror copy (outflow); j, k interior
__global__ void __launch_bounds__(kBlock) k_ghost_x(Grid g, int nprim, double *prim, double *T, double *E)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long per = (long long)g.bc * g.nj * g.nk;
    if (t >= 2 * per)
        return;
    int side = (int)(t / per);
    long long r = t % per;
    int k = (int)(r % g.nk) + g.bc;
    r /= g.nk;
    int j = (int)(r % g.nj) + g.bc;
    int ii = (int)(r / g.nj);
    int i, is;
    if (side == 0) {
        i = ii;
        is = 2 * g.bc - 1 - i;
    } else {
        i = g.ni + g.bc + ii;
        is = 2 * (g.ni + g.bc) - 1 - i;
    }
    copy_point(g, nprim, prim, T, E, gidx(g, i, j, k), gidx(g, is, j, k), -1);
}

// y faces: front symmetry (v -> -v, thermo recomputed), back copy; all i, k interior
template <int NS, class MT>
__global__ void __launch_bounds__(kBlock)
    k_ghost_y(Grid g, const DevMech *__restrict__ mp, double *prim, double *T, double *E)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long per = (long long)g.NI * g.bc * g.nk;
    if (t >= 2 * per)
        return;
    int side = (int)(t / per);
    long long r = t % per;
    int k = (int)(r % g.nk) + g.bc;
    r /= g.nk;
    int jj = (int)(r % g.bc);
    int i = (int)(r / g.bc);
    const long long N = g.N;
    if (side == 1) {
        int j = g.nj + g.bc + jj;
        int js = 2 * (g.nj + g.bc) - 1 - j;
        copy_point(g, 7 + NS, prim, T, E, gidx(g, i, j, k), gidx(g, i, js, k), -1);
        return;
    }
    const auto &m = mech_ref<MT>(mp);
    int j = jj, js = 2 * g.bc - 1 - j;
    long long d = gidx(g, i, j, k), s = gidx(g, i, js, k);
    double Y[NS];
    prim[2 * N + d] = prim[2 * N + s];
    prim[3 * N + d] = -prim[3 * N + s];
    prim[4 * N + d] = prim[4 * N + s];
    double P = prim[0 * N + s];
    prim[0 * N + d] = P;
    double Tt = T[s];
    T[d] = Tt;
#pragma unroll
    for (int q = 0; q < NS; q++) {
        Y[q] = prim[(7 + q) * N + s];
        prim[(7 + q) * N + d] = Y[q];
    }
    TPow tp = tpow(Tt);
    double sR = 0.0, sCp = 0.0, sH = 0.0, sE = 0.0;
#pragma unroll
    for (int q = 0; q < NS; q++)
        sR += Y[q] * m.ri[q];
    by_range(m, Tt, [&](const auto &a) {
#pragma unroll
        for (int q = 0; q < NS; q++) {
            double cp = getcpi(tp, m.ri[q], a, q, NS);
            double h = gethi(tp, m.ri[q], a, q, NS);
            double e = h - m.ri[q] * Tt;
            sCp += cp * Y[q];
            sH += Y[q] * h;
            sE += Y[q] * e;
        }
    });
    double Rgas = sR * 1000;
    double D = P / (Tt * Rgas);
    double Cp = sCp * 1000;
    prim[1 * N + d] = D;
    prim[5 * N + d] = sH * 1000;
    E[d] = sE * 1000;
    prim[6 * N + d] = Cp / (Cp - Rgas);
}

// z faces: w -> -w mirror on both sides; all i, j
__global__ void __launch_bounds__(kBlock) k_ghost_z(Grid g, int nprim, double *prim, double *T, double *E)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long per = (long long)g.NI * g.NJ * g.bc;
    if (t >= 2 * per)
        return;
    int side = (int)(t / per);
    long long r = t % per;
    int kk = (int)(r % g.bc);
    r /= g.bc;
    int j = (int)(r % g.NJ);
    int i = (int)(r / g.NJ);
    int k, ks;
    if (side == 0) {
        k = kk;
        ks = 2 * g.bc - 1 - k;
    } else {
        k = g.nk + g.bc + kk;
        ks = 2 * (g.nk + g.bc) - 1 - k;
    }
    copy_point(g, nprim, prim, T, E, gidx(g, i, j, k), gidx(g, i, j, ks), 4);
}

// ---------------------------------------------------------------------------
// Fused flux + conservative update.  MODE 0: EE (CS += RHS); 1..3: SSP-RK3
// stages (1 also saves CS into CS0).
//
// A block owns kTJ rows (j) of kTK = 64 cells (k, one wave per row) and
// marches over kMarch planes (i).  Every face flux is computed once:
//   x faces: by the owning thread, the plus face of plane i is kept in
//            registers as the minus face of plane i+1;
//   y faces: kTJ+1 rows of 64 faces, one per thread (the extra wave computes
//            the faces below the tile), exchanged through LDS;
//   z faces: kTJ rows of 65 faces on the first kTJ*65 threads, through LDS.
// RHS = -dt/dx (F+ - F-) - dt/dy (G+ - G-) - dt/dz (Q+ - Q-) in that order
// (TimeAdv::EE, src/TimeAdv.cpp:51-54).
// ---------------------------------------------------------------------------
constexpr int kTJ = 4;
constexpr int kTK = 64;
constexpr int kFluxThreads = 64 * (kTJ + 1);
constexpr int kMarch = 16;
#ifndef EU3D_FLUX_WAVES_PER_EU
#define EU3D_FLUX_WAVES_PER_EU 2
#endif

template <int NS, int MODE>
__global__ void __launch_bounds__(kFluxThreads) __attribute__((amdgpu_waves_per_eu(EU3D_FLUX_WAVES_PER_EU)))
    k_flux_tile(Grid g, const double *__restrict__ prim, double *__restrict__ CS, double *__restrict__ CS0,
                const double *__restrict__ xn, const double *__restrict__ yn, const double *__restrict__ zn,
                double dt)
{
    constexpr int NV = NS + 4;
    constexpr int NYF = (kTJ + 1) * kTK; // y faces per plane
    constexpr int NZF = kTJ * (kTK + 1); // z faces per plane
    static_assert(NZF <= NYF, "LDS face buffer sized for the y faces");
    __shared__ double face[NV * NYF];
    const int t = threadIdx.x;
    const long long N = g.N;
    const int k0 = g.bc + blockIdx.x * kTK;
    const int j0 = g.bc + blockIdx.y * kTJ;
    const int i0 = g.bc + blockIdx.z * kMarch;
    const int i1 = min(i0 + kMarch, g.ni + g.bc);
    const int kend = g.nk + g.bc, jend = g.nj + g.bc;
    // the cell this thread owns
    const int r = t >> 6, l = t & 63;
    const int jc = j0 + r, kc = k0 + l;
    const bool own = (t < kTJ * kTK) && (jc < jend) && (kc < kend);
    // the y face this thread computes: between j = yj and yj+1 at k = yk
    const int yj = j0 - 1 + (t >> 6), yk = k0 + (t & 63);
    const bool yact = (yj < jend) && (yk < kend);
    // the z face this thread computes: between k = zk and zk+1 in row zj
    const int zr = t / (kTK + 1), zf = t - zr * (kTK + 1);
    const int zj = j0 + zr, zk = k0 - 1 + zf;
    const bool zact = (t < NZF) && (zj < jend) && (zk < kend);

    double rhs[NV], Fxm[NV];
    if (own)
        face_flux<NS>(0, prim, N, gidx(g, i0 - 1, jc, kc), g.sI, Fxm);
    for (int i = i0; i < i1; i++) {
        if (own) {
            double Fxp[NV];
            face_flux<NS>(0, prim, N, gidx(g, i, jc, kc), g.sI, Fxp);
            double cx = -dt / (xn[i] - xn[i - 1]);
#pragma unroll
            for (int q = 0; q < NV; q++) {
                rhs[q] = cx * (Fxp[q] - Fxm[q]);
                Fxm[q] = Fxp[q];
            }
        }
        if (yact) {
            double F[NV];
            face_flux<NS>(1, prim, N, gidx(g, i, yj, yk), g.sJ, F);
#pragma unroll
            for (int q = 0; q < NV; q++)
                face[q * NYF + t] = F[q];
        }
        __syncthreads();
        if (own) {
            double cy = dt / (yn[jc] - yn[jc - 1]);
#pragma unroll
            for (int q = 0; q < NV; q++)
                rhs[q] = rhs[q] - cy * (face[q * NYF + (r + 1) * kTK + l] - face[q * NYF + r * kTK + l]);
        }
        __syncthreads();
        if (zact) {
            double F[NV];
            face_flux<NS>(2, prim, N, gidx(g, i, zj, zk), 1, F);
#pragma unroll
            for (int q = 0; q < NV; q++)
                face[q * NZF + t] = F[q];
        }
        __syncthreads();
        if (own) {
            double cz = dt / (zn[kc] - zn[kc - 1]);
            const long long c = gidx(g, i, jc, kc);
#pragma unroll
            for (int q = 0; q < NV; q++) {
                double rq = rhs[q] - cz * (face[q * NZF + r * (kTK + 1) + l + 1] - face[q * NZF + r * (kTK + 1) + l]);
                double u = CS[q * N + c];
                double v;
                if (MODE == 0) {
                    v = u + rq;
                } else if (MODE == 1) {
                    CS0[q * N + c] = u;
                    v = u + rq;
                } else if (MODE == 2) {
                    v = 0.75 * CS0[q * N + c] + 0.25 * (u + rq);
                } else {
                    v = (1.0 / 3.0) * CS0[q * N + c] + (2.0 / 3.0) * (u + rq);
                }
                CS[q * N + c] = v;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// state recovery (Update_after_CS) over the interior
// ---------------------------------------------------------------------------
template <int NS>
__device__ __forceinline__ void store_state(long long N, long long c, const CellState<NS> &st, double *prim,
                                            double *T, double *E, double *C)
{
    prim[0 * N + c] = st.P;
    prim[1 * N + c] = st.D;
    prim[2 * N + c] = st.U;
    prim[3 * N + c] = st.V;
    prim[4 * N + c] = st.W;
    prim[5 * N + c] = st.H;
    prim[6 * N + c] = st.G;
#pragma unroll
    for (int s = 0; s < NS; s++)
        prim[(7 + s) * N + c] = st.Y[s];
    T[c] = st.T;
    E[c] = st.E;
    C[c] = st.C;
}

template <int NS, class MT>
__global__ void __launch_bounds__(kBlock)
    k_recover(Grid g, const DevMech *__restrict__ mp, const double *__restrict__ CS, double *__restrict__ prim,
              double *__restrict__ T, double *__restrict__ E, double *__restrict__ C)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!interior_cell(g, t, i, j, k))
        return;
    const long long N = g.N, c = gidx(g, i, j, k);
    double cs[NS + 4];
#pragma unroll
    for (int q = 0; q < NS + 4; q++)
        cs[q] = CS[q * N + c];
    CellState<NS> st;
    recover<NS>(mech_ref<MT>(mp), cs, T[c], E[c], st);
    store_state<NS>(N, c, st, prim, T, E, C);
}

// ---------------------------------------------------------------------------
// CFL: min over the interior of min(dx/(|u|+c), dy/(|v|+c)); NaN ignored as
// by the reference's `if (temp < dt_conv)`.  Positive doubles order like
// their bit patterns, so the minimum is an order-free atomicMin.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
    k_cfl(Grid g, const double *__restrict__ prim, const double *__restrict__ C, double dx, double dy,
          unsigned long long *out)
{
    __shared__ double red[kBlock / 64];
    double best = 1e10;
    long long n_int = (long long)g.ni * g.nj * g.nk;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n_int;
         t += (long long)gridDim.x * blockDim.x) {
        int i, j, k;
        interior_cell(g, t, i, j, k);
        long long c = gidx(g, i, j, k);
        double cc = C[c];
        double a = dx / (fabs(prim[2 * g.N + c]) + cc);
        double b = dy / (fabs(prim[3 * g.N + c]) + cc);
        double temp = (b < a) ? b : a;
        if (temp < best)
            best = temp;
    }
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(best, off, 64);
        if (o < best)
            best = o;
    }
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        red[wid] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < kBlock / 64; w++)
            if (red[w] < b)
                b = red[w];
        atomicMin(out, (unsigned long long)__double_as_longlong(b));
    }
}

// ---------------------------------------------------------------------------
// chemistry
// ---------------------------------------------------------------------------
__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

__global__ void k_chem_reset(ChemCounters *c)
{
    c->queue_count = 0;
    c->fetch = 0;
    c->invalid = 0;
    c->nonfinite = 0;
    c->stiff = 0;
    c->substeps = 0;
    c->max_nchem = 0;
}

template <int NS>
struct ChemCell {
    double Di[NS]; // SI partial densities (= CS[0..NS))
    double T, E, U, V, W;
};

template <int NS>
__device__ __forceinline__ void chem_load(const Grid &g, long long c, const double *CS, const double *prim,
                                          const double *T, const double *E, ChemCell<NS> &x)
{
    const long long N = g.N;
#pragma unroll
    for (int s = 0; s < NS; s++)
        x.Di[s] = CS[s * N + c];
    x.T = T[c];
    x.E = E[c];
    x.U = prim[2 * N + c];
    x.V = prim[3 * N + c];
    x.W = prim[4 * N + c];
}

// One sub-step: trapezoid update (src/Reaction.cpp:781-805) with rates R
// already evaluated at the current state, then Explicit(step) +
// Update_after_CS(step) (src/Flowfield.cpp:1522-1536, 1620-1676).
template <int NS, class MT>
__device__ __forceinline__ void chem_substep(const MT &m, ChemCell<NS> &x, const Rates<NS> &R,
                                             const double (&Mc)[NS], double dt, double (&cms)[NS],
                                             double (&cs)[NS + 4], CellState<NS> &st)
{
#pragma unroll
    for (int s = 0; s < NS; s++) {
        cms[s] = R.W[s] * 1e3;
        double Di = x.Di[s] * 1e-3;
        double Qs = m.mw[s] * R.Q[s];
        double Ps;
        if (Mc[s] == 0)
            Ps = 0.0;
        else
            Ps = R.P[s] / Mc[s];
        x.Di[s] = ((1.0 - dt / 2.0 * Ps) * Di + dt * Qs) / (1.0 + dt / 2.0 * Ps) * 1e3;
    }
    double D = 0.0;
#pragma unroll
    for (int s = 0; s < NS; s++) {
        cs[s] = x.Di[s];
        D += cs[s];
    }
    cs[NS] = x.U * D;
    cs[NS + 1] = x.V * D;
    cs[NS + 2] = x.W * D;
    cs[NS + 3] = D * x.E + 0.5 * ((x.U * x.U) + (x.V * x.V) + (x.W * x.W)) * D;
    recover<NS>(m, cs, x.T, x.E, st);
    x.T = st.T;
    x.E = st.E;
    x.U = st.U;
    x.V = st.V;
    x.W = st.W;
}

template <int NS, class MT>
__device__ __forceinline__ void chem_mc(const MT &m, const ChemCell<NS> &x, double (&Mc)[NS])
{
    // Mc = Di/Mw*1000 (Update_after_CS) then *1e-6 (Reaction::Trapezoid)
#pragma unroll
    for (int s = 0; s < NS; s++)
        Mc[s] = x.Di[s] / m.mw[s] * 1000 * 1e-6;
}

template <int NS>
__device__ __forceinline__ void chem_store(long long N, long long c, const double (&cs)[NS + 4],
                                           const CellState<NS> &st, const double (&cms)[NS], int keep_cms,
                                           double *CS, double *CMS, double *prim, double *T, double *E, double *C)
{
#pragma unroll
    for (int q = 0; q < NS + 4; q++)
        CS[q * N + c] = cs[q];
    if (keep_cms) {
#pragma unroll
        for (int s = 0; s < NS; s++)
            CMS[s * N + c] = cms[s];
    }
    store_state<NS>(N, c, st, prim, T, E, C);
}

#ifndef EU3D_CHEM_WAVES_PER_EU
#define EU3D_CHEM_WAVES_PER_EU 2
#endif

__device__ __forceinline__ unsigned long long lanemask_lt()
{
    unsigned int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// One sub-step of cell c from the state in memory, written back to memory.
// FIRST: the rates also give Nchem (TrapezoidPrediction, src/Reaction.cpp:
// 865-942), and the sub-step uses them (the reference evaluates the same
// rates twice on identical inputs).  Returns Nchem (0: invalid, no update).
// The rest of the cell state is read after the rates so that it is not held
// in registers through them.
template <int NS, bool FIRST, class MT>
__device__ __forceinline__ int chem_step_cell(const Grid &g, const MT &m, long long c, double dtm, int n,
                                              double *CS, double *prim, double *T, double *E, double *C,
                                              double *CMS, int keep_cms)
{
    const long long N = g.N;
    const double Tc = T[c];
    double Mc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++)
        Mc[s] = CS[s * N + c] / m.mw[s] * 1000 * 1e-6; // Update_after_CS, then Reaction::Trapezoid
    Rates<NS> R;
    rates<NS, true>(m, Tc, Mc, R);
    long long cl = c;
    asm volatile("" : "+v"(cl) : "v"(R.W[0]));
    ChemCell<NS> x;
    chem_load<NS>(g, cl, CS, prim, T, E, x);
    if (FIRST) {
        double Dcgs[NS], Y[NS];
#pragma unroll
        for (int s = 0; s < NS; s++) {
            Dcgs[s] = x.Di[s] * 1e-3;
            Y[s] = prim[(7 + s) * N + cl];
        }
        n = predict_nchem<NS>(R.W, Dcgs, Y, dtm);
        if (n < 1)
            return n;
    }
    double cms[NS], cs[NS + 4];
    CellState<NS> st;
    chem_substep<NS>(m, x, R, Mc, dtm / n, cms, cs, st);
    chem_store<NS>(N, cl, cs, st, cms, keep_cms, CS, CMS, prim, T, E, C);
    return n;
}

// First chemistry pass over every interior cell: rates at the post-advection
// state, Nchem (TrapezoidPrediction), and sub-step 0 with the same rates (the
// reference evaluates them twice on identical inputs).  Cells with Nchem > 1
// are appended to the stiff queue.
template <int NS, class MT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(EU3D_CHEM_WAVES_PER_EU)))
    k_chem_first(Grid g, const DevMech *__restrict__ mp, double dtm, double *__restrict__ CS,
                 double *__restrict__ prim, double *__restrict__ T, double *__restrict__ E, double *__restrict__ C,
                 double *__restrict__ CMS, int *__restrict__ nchem, int *__restrict__ queue, ChemCounters *cnt,
                 int compact, int keep_cms)
{
    const auto &m = mech_ref<MT>(mp);
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i = 0, j = 0, k = 0;
    bool active = interior_cell(g, t, i, j, k);
    long long c = active ? gidx(g, i, j, k) : 0;
    int n = 0;
    bool stiff = false;
    if (active) {
        n = chem_step_cell<NS, true>(g, m, c, dtm, 0, CS, prim, T, E, C, CMS, keep_cms);
        nchem[c] = n;
        if (n >= 1)
            stiff = n > 1;
        else
            atomicAdd(&cnt->invalid, 1ull);
    }
    // wave-aggregated append of stiff cells
    unsigned long long mask = __ballot(stiff);
    if (mask) {
        unsigned int base = 0;
        int leader = __ffsll((long long)mask) - 1;
        if ((int)(threadIdx.x & 63) == leader)
            base = atomicAdd(&cnt->queue_count, (unsigned int)__popcll(mask));
        base = __shfl(base, leader, 64);
        if (stiff && compact)
            queue[base + __popcll(mask & lanemask_lt())] = (int)c;
    }
    // health: substeps and max Nchem
    long long sub = active ? n : 0;
    for (int off = 32; off > 0; off >>= 1)
        sub += __shfl_xor(sub, off, 64);
    int mx = n;
    for (int off = 32; off > 0; off >>= 1)
        mx = max(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) {
        if (sub)
            atomicAdd(&cnt->substeps, (unsigned long long)sub);
        atomicMax(&cnt->max_nchem, mx);
    }
}

// Remaining sub-steps 1..Nchem-1 of one cell.
template <int NS, class MT>
__device__ __forceinline__ void chem_rest(const Grid &g, const MT &m, long long c, int n, double dtm,
                                          double *CS, double *prim, double *T, double *E, double *C, double *CMS,
                                          int keep_cms)
{
    for (int step = 1; step < n; step++)
        chem_step_cell<NS, false>(g, m, c, dtm, n, CS, prim, T, E, C, CMS, keep_cms);
}

// Compacted: waves fetch 64 queued cells at a time until the queue is drained.
template <int NS, class MT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(EU3D_CHEM_WAVES_PER_EU)))
    k_chem_queue(Grid g, const DevMech *__restrict__ mp, double dtm, double *__restrict__ CS,
                 double *__restrict__ prim, double *__restrict__ T, double *__restrict__ E, double *__restrict__ C,
                 double *__restrict__ CMS, const int *__restrict__ nchem, const int *__restrict__ queue,
                 ChemCounters *cnt, int keep_cms)
{
    const unsigned int total = cnt->queue_count;
    const int lane = threadIdx.x & 63;
    for (;;) {
        unsigned int base = 0;
        if (lane == 0)
            base = atomicAdd(&cnt->fetch, 64u);
        base = __shfl(base, 0, 64);
        if (base >= total)
            break;
        unsigned int q = base + lane;
        if (q < total) {
            long long c = queue[q];
            chem_rest<NS>(g, mech_ref<MT>(mp), c, nchem[c], dtm, CS, prim, T, E, C, CMS, keep_cms);
        }
    }
}

// Uncompacted: every interior cell runs its own remaining sub-steps.
template <int NS, class MT>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(EU3D_CHEM_WAVES_PER_EU)))
    k_chem_all(Grid g, const DevMech *__restrict__ mp, double dtm, double *__restrict__ CS,
               double *__restrict__ prim, double *__restrict__ T, double *__restrict__ E, double *__restrict__ C,
               double *__restrict__ CMS, const int *__restrict__ nchem, int keep_cms)
{
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k;
    if (!interior_cell(g, t, i, j, k))
        return;
    long long c = gidx(g, i, j, k);
    int n = nchem[c];
    if (n > 1)
        chem_rest<NS>(g, mech_ref<MT>(mp), c, n, dtm, CS, prim, T, E, C, CMS, keep_cms);
}

// ---------------------------------------------------------------------------
// probes
// ---------------------------------------------------------------------------
template <int NS, class MT>
__global__ void k_probe_thermo(const DevMech *__restrict__ mp, const double *Tin, int n, double R, double *cp,
                               double *h, double *s)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * NS)
        return;
    int it = t / NS, sp = t % NS;
    double T = Tin[it];
    TPow tp = tpow(T);
    const auto &m = mech_ref<MT>(mp);
    const CoefPtr a{T < EU3D_THERMO_SWITCH_T ? m.c0 : m.c1};
    cp[t] = getcpi(tp, R, a, sp, NS);
    h[t] = gethi(tp, R, a, sp, NS);
    s[t] = getsi(tp, R, a, sp, NS);
}

template <int NS>
__global__ void k_probe_flux(Grid g, int dir, const double *__restrict__ prim, double *__restrict__ out)
{
    constexpr int NV = NS + 4;
    long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    long long ni = g.ni + 1, nj = g.nj + 1, nk = g.nk + 1; // [bc-1, n+bc)
    if (t >= ni * nj * nk)
        return;
    int k = (int)(t % nk) + g.bc - 1;
    long long r = t / nk;
    int j = (int)(r % nj) + g.bc - 1;
    int i = (int)(r / nj) + g.bc - 1;
    long long a = gidx(g, i, j, k);
    long long st = dir == 0 ? g.sI : (dir == 1 ? g.sJ : 1);
    double F[NV];
    face_flux<NS>(dir, prim, g.N, a, st, F);
#pragma unroll
    for (int q = 0; q < NV; q++)
        out[q * g.N + a] = F[q];
}

template <int NS, class MT>
__global__ void k_probe_rates(const DevMech *__restrict__ mp, const double *Tin, const double *Mcin, int n,
                              double *wi)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    double Mc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++)
        Mc[s] = Mcin[t * NS + s];
    Rates<NS> R;
    rates<NS, false>(mech_ref<MT>(mp), Tin[t], Mc, R);
#pragma unroll
    for (int s = 0; s < NS; s++)
        wi[t * NS + s] = R.W[s];
}

// Cp of every point for downloads (Update_after_CS, src/Flowfield.cpp:1609-1611)
template <int NS, class MT>
__global__ void k_derive_cp(Grid g, const DevMech *__restrict__ mp, const double *__restrict__ prim,
                            const double *__restrict__ T, double *__restrict__ cp_out)
{
    long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (c >= g.N)
        return;
    const auto &m = mech_ref<MT>(mp);
    double Tc = T[c];
    TPow tp = tpow(Tc);
    double scp = 0.0;
    by_range(m, Tc, [&](const auto &a) {
#pragma unroll
        for (int s = 0; s < NS; s++)
            scp += getcpi(tp, m.ri[s], a, s, NS) * prim[(7 + s) * g.N + c];
    });
    cp_out[c] = scp * 1000;
}

// ---------------------------------------------------------------------------
// dispatch over the compiled species counts
// ---------------------------------------------------------------------------
#define EU3D_DISPATCH_NS(ns, F, ...)                                                              \
    switch (ns) {                                                                                 \
    case 9:                                                                                       \
        if (ctx->generated)                                                                       \
            F<9, MechH2JA919>(__VA_ARGS__);                                                       \
        else                                                                                      \
            F<9, DevMech>(__VA_ARGS__);                                                           \
        break;                                                                                    \
    default:                                                                                      \
        FAIL(EU3D_ERR_ARG, "species count not compiled in (supported: 9)");                       \
    }

template <int NS, class MT>
static void launch_flux(eu3d_ctx *ctx, int mode, double dt)
{
    const Grid &g = ctx->g;
    dim3 grid((g.nk + kTK - 1) / kTK, (g.nj + kTJ - 1) / kTJ, (g.ni + kMarch - 1) / kMarch);
    switch (mode) {
    case 0:
        k_flux_tile<NS, 0><<<grid, kFluxThreads, 0, ctx->stream>>>(g, ctx->prim, ctx->CS, ctx->CS0, ctx->x, ctx->y,
                                                                    ctx->z, dt);
        break;
    case 1:
        k_flux_tile<NS, 1><<<grid, kFluxThreads, 0, ctx->stream>>>(g, ctx->prim, ctx->CS, ctx->CS0, ctx->x, ctx->y,
                                                                    ctx->z, dt);
        break;
    case 2:
        k_flux_tile<NS, 2><<<grid, kFluxThreads, 0, ctx->stream>>>(g, ctx->prim, ctx->CS, ctx->CS0, ctx->x, ctx->y,
                                                                    ctx->z, dt);
        break;
    default:
        k_flux_tile<NS, 3><<<grid, kFluxThreads, 0, ctx->stream>>>(g, ctx->prim, ctx->CS, ctx->CS0, ctx->x, ctx->y,
                                                                    ctx->z, dt);
        break;
    }
}

template <int NS, class MT>
static void launch_recover(eu3d_ctx *ctx)
{
    k_recover<NS, MT><<<blocks_for(ctx->n_int), kBlock, 0, ctx->stream>>>(ctx->g, ctx->mech, ctx->CS, ctx->prim, ctx->T,
                                                                       ctx->E, ctx->C);
}

template <int NS, class MT>
static void launch_ghost_y(eu3d_ctx *ctx)
{
    long long n = 2LL * ctx->g.NI * ctx->g.bc * ctx->g.nk;
    k_ghost_y<NS, MT><<<blocks_for(n), kBlock, 0, ctx->stream>>>(ctx->g, ctx->mech, ctx->prim, ctx->T, ctx->E);
}

template <int NS, class MT>
static void launch_chem(eu3d_ctx *ctx, double dt, int flags)
{
    const Grid &g = ctx->g;
    int compact = (flags & EU3D_CHEM_COMPACT) ? 1 : 0;
    int keep = (flags & EU3D_CHEM_KEEP_CMS) ? 1 : 0;
    k_chem_reset<<<1, 1, 0, ctx->stream>>>(ctx->cnt);
    k_chem_first<NS, MT><<<blocks_for(ctx->n_int), kBlock, 0, ctx->stream>>>(
        g, ctx->mech, dt, ctx->CS, ctx->prim, ctx->T, ctx->E, ctx->C, ctx->CMS, ctx->nchem, ctx->queue, ctx->cnt,
        compact, keep);
    if (compact) {
        k_chem_queue<NS, MT><<<ctx->ncu * 4, kBlock, 0, ctx->stream>>>(g, ctx->mech, dt, ctx->CS, ctx->prim, ctx->T, ctx->E,
                                                              ctx->C, ctx->CMS, ctx->nchem, ctx->queue, ctx->cnt,
                                                              keep);
    } else {
        k_chem_all<NS, MT><<<blocks_for(ctx->n_int), kBlock, 0, ctx->stream>>>(g, ctx->mech, dt, ctx->CS, ctx->prim,
                                                                           ctx->T, ctx->E, ctx->C, ctx->CMS,
                                                                           ctx->nchem, keep);
    }
}

template <int NS, class MT>
static void launch_probe_thermo(eu3d_ctx *ctx, const double *T, int n, double R, double *cp, double *h, double *s)
{
    k_probe_thermo<NS, MT><<<blocks_for((long long)n * NS), kBlock, 0, ctx->stream>>>(ctx->mech, T, n, R, cp, h, s);
}

template <int NS, class MT>
static void launch_probe_flux(eu3d_ctx *ctx, int dir, double *out)
{
    long long n = (long long)(ctx->g.ni + 1) * (ctx->g.nj + 1) * (ctx->g.nk + 1);
    k_probe_flux<NS><<<blocks_for(n), kBlock, 0, ctx->stream>>>(ctx->g, dir, ctx->prim, out);
}

template <int NS, class MT>
static void launch_probe_rates(eu3d_ctx *ctx, const double *T, const double *Mc, int n, double *wi)
{
    k_probe_rates<NS, MT><<<blocks_for(n), kBlock, 0, ctx->stream>>>(ctx->mech, T, Mc, n, wi);
}

template <int NS, class MT>
static void launch_derive_cp(eu3d_ctx *ctx, double *out)
{
    k_derive_cp<NS, MT><<<blocks_for(ctx->g.N), kBlock, 0, ctx->stream>>>(ctx->g, ctx->mech, ctx->prim, ctx->T, out);
}

// True when the run-time tables equal the generated ones bit for bit.
template <class G>
static bool tables_equal(const DevMech &m)
{
    auto same = [](const double *a, const double *b, int n) { return std::memcmp(a, b, sizeof(double) * n) == 0; };
    double R = G::R, Ru = G::Ru, P0 = G::P0;
    return m.ns == G::ns && m.nr == G::nr && same(&m.R, &R, 1) && same(&m.Ru, &Ru, 1) && same(&m.P0, &P0, 1) &&
           same(m.mw, G::mw, kMaxNS) && same(m.ri, G::ri, kMaxNS) && same(m.c0, G::c0, 9 * kMaxNS) &&
           same(m.c1, G::c1, 9 * kMaxNS) && same(m.af, G::af, kMaxNR) && same(m.bf, G::bf, kMaxNR) &&
           same(m.eaf, G::eaf, kMaxNR) && same(m.nf, G::nf, kMaxNS * kMaxNR) &&
           same(m.nb, G::nb, kMaxNS * kMaxNR) && same(m.tb, G::tb, kMaxNS * kMaxNR) &&
           same(m.p0x, G::p0x, kMaxNR) && same(m.mxr, G::mxr, kMaxNR);
}

static bool mech_matches_generated(const DevMech &m) { return tables_equal<MechH2JA919>(m); }

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int eu3d_set_option(eu3d_ctx *ctx, int option, int value)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (option == EU3D_OPT_GENERIC_MECHANISM) {
        ctx->generated = ctx->generated_available && value == 0;
        return EU3D_OK;
    }
    FAIL(EU3D_ERR_ARG, "unknown option");
}

int eu3d_get_option(eu3d_ctx *ctx, int option, int *value)
{
    if (!ctx || !value)
        return EU3D_ERR_ARG;
    if (option == EU3D_OPT_GENERIC_MECHANISM) {
        *value = ctx->generated ? 0 : 1;
        return EU3D_OK;
    }
    if (option == EU3D_OPT_GENERATED_AVAILABLE) {
        *value = ctx->generated_available ? 1 : 0;
        return EU3D_OK;
    }
    FAIL(EU3D_ERR_ARG, "unknown option");
}

const char *eu3d_last_error(const eu3d_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int eu3d_destroy(eu3d_ctx *ctx)
{
    if (!ctx)
        return EU3D_OK;
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    resolve_timing(ctx);
    for (auto e : ctx->free_events)
        (void)hipEventDestroy(e);
    void *ptrs[] = {ctx->mech, ctx->x,   ctx->y,   ctx->z,       ctx->prim,  ctx->T,     ctx->E,  ctx->C,
                    ctx->CS,   ctx->CS0, ctx->CMS, ctx->scratch, ctx->nchem, ctx->queue, ctx->cnt};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return EU3D_OK;
}

int eu3d_create(const eu3d_mesh *mesh, const eu3d_mechanism *mc, eu3d_ctx **out)
{
    if (!out)
        return EU3D_ERR_ARG;
    *out = nullptr;
    if (!mesh || !mc)
        return EU3D_ERR_ARG;
    eu3d_ctx *ctx = new eu3d_ctx();
    *out = ctx;
    if (mesh->ni < 3 || mesh->nj < 3 || mesh->nk < 3 || mesh->bc < 2 || mesh->ni < mesh->bc ||
        mesh->nj < mesh->bc || mesh->nk < mesh->bc)
        FAIL(EU3D_ERR_ARG, "grid must have >= max(3, bc) interior nodes per direction and bc >= 2");
    if (mc->ns < 1 || mc->ns > EU3D_MAX_NS || mc->nr < 1 || mc->nr > EU3D_MAX_NR)
        FAIL(EU3D_ERR_ARG, "NS must be in [1,16] and NR in [1,24] (include/ReactionCuda.cuh:4-5)");
    if (mc->ns != 9)
        FAIL(EU3D_ERR_ARG, "species count not compiled in (supported: 9)");
    Grid &g = ctx->g;
    g.ni = mesh->ni;
    g.nj = mesh->nj;
    g.nk = mesh->nk;
    g.bc = mesh->bc;
    g.NI = g.ni + 2 * g.bc;
    g.NJ = g.nj + 2 * g.bc;
    g.NK = g.nk + 2 * g.bc;
    g.N = (long long)g.NI * g.NJ * g.NK;
    g.sJ = g.NK;
    g.sI = (long long)g.NJ * g.NK;
    ctx->ns = mc->ns;
    ctx->nr = mc->nr;
    ctx->nv = mc->ns + 4;
    ctx->nprim = 7 + mc->ns;
    ctx->n_int = (long long)g.ni * g.nj * g.nk;

    DevMech &m = ctx->hmech;
    std::memset(&m, 0, sizeof m);
    m.ns = mc->ns;
    m.nr = mc->nr;
    m.R = mc->R;
    m.Ru = mc->Ru;
    m.P0 = mc->P0;
    int ns = mc->ns, nr = mc->nr;
    for (int s = 0; s < ns; s++) {
        m.mw[s] = mc->mw[s];
        m.ri[s] = mc->ri[s];
    }
    for (int q = 0; q < 9 * ns; q++) {
        m.c0[q] = mc->coeff0[q];
        m.c1[q] = mc->coeff1[q];
    }
    for (int r = 0; r < nr; r++) {
        m.af[r] = mc->af[r];
        m.bf[r] = mc->bf[r];
        m.eaf[r] = mc->eaf[r];
        double Xr = 0.0;
        for (int s = 0; s < ns; s++)
            Xr += (mc->stoi_b[s * nr + r] - mc->stoi_f[s * nr + r]);
        m.p0x[r] = std::pow((m.P0 * 1e-6), Xr);
        m.mxr[r] = -Xr;
    }
    for (int q = 0; q < ns * nr; q++) {
        m.nf[q] = mc->stoi_f[q];
        m.nb[q] = mc->stoi_b[q];
        m.tb[q] = mc->react_tb[q];
    }

    {
        int dev = 0;
        CK(hipGetDevice(&dev));
        CK(hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, dev));
    }
    ctx->generated = mech_matches_generated(m);
    ctx->generated_available = ctx->generated;
    CK(hipStreamCreate(&ctx->stream));
    CK(hipMalloc(&ctx->mech, sizeof(DevMech)));
    CK(hipMemcpy(ctx->mech, &m, sizeof(DevMech), hipMemcpyHostToDevice));
    CK(hipMalloc(&ctx->x, sizeof(double) * g.NI));
    CK(hipMalloc(&ctx->y, sizeof(double) * g.NJ));
    CK(hipMalloc(&ctx->z, sizeof(double) * g.NK));
    CK(hipMemcpy(ctx->x, mesh->xnode, sizeof(double) * g.NI, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->y, mesh->ynode, sizeof(double) * g.NJ, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->z, mesh->znode, sizeof(double) * g.NK, hipMemcpyHostToDevice));
    size_t N = (size_t)g.N;
    CK(hipMalloc(&ctx->prim, sizeof(double) * N * ctx->nprim));
    CK(hipMalloc(&ctx->T, sizeof(double) * N));
    CK(hipMalloc(&ctx->E, sizeof(double) * N));
    CK(hipMalloc(&ctx->C, sizeof(double) * N));
    CK(hipMalloc(&ctx->CS, sizeof(double) * N * ctx->nv));
    CK(hipMalloc(&ctx->CS0, sizeof(double) * N * ctx->nv));
    CK(hipMalloc(&ctx->CMS, sizeof(double) * N * ns));
    CK(hipMalloc(&ctx->scratch, sizeof(double) * N));
    CK(hipMalloc(&ctx->nchem, sizeof(int) * N));
    CK(hipMalloc(&ctx->queue, sizeof(int) * (size_t)ctx->n_int));
    CK(hipMalloc(&ctx->cnt, sizeof(ChemCounters)));
    CK(hipMemset(ctx->cnt, 0, sizeof(ChemCounters)));
    CK(hipMemset(ctx->CMS, 0, sizeof(double) * N * ns));
    CK(hipMemset(ctx->CS0, 0, sizeof(double) * N * ctx->nv));
    std::vector<int> ones(N, 1); // Nchem.Fill(1) (src/Reaction.cpp:102)
    CK(hipMemcpy(ctx->nchem, ones.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    ctx->has_state = false;
    return EU3D_OK;
}

int eu3d_upload_state(eu3d_ctx *ctx, const eu3d_state *h)
{
    if (!ctx || !h)
        return EU3D_ERR_ARG;
    const size_t N = (size_t)ctx->g.N;
    const int ns = ctx->ns, nv = ctx->nv;
    if (!h->P || !h->D || !h->U || !h->V || !h->W || !h->H || !h->Gamma || !h->Yi || !h->T || !h->E || !h->CS)
        FAIL(EU3D_ERR_ARG, "upload needs P D U V W H Gamma Yi T E CS");
    std::vector<double> buf(N * std::max(nv, 7 + ns));
    const double *base[7] = {h->P, h->D, h->U, h->V, h->W, h->H, h->Gamma};
    for (int f = 0; f < 7; f++)
        std::memcpy(buf.data() + f * N, base[f], sizeof(double) * N);
    for (size_t c = 0; c < N; c++)
        for (int s = 0; s < ns; s++)
            buf[(7 + s) * N + c] = h->Yi[c * ns + s];
    CK(hipMemcpy(ctx->prim, buf.data(), sizeof(double) * N * (7 + ns), hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->T, h->T, sizeof(double) * N, hipMemcpyHostToDevice));
    CK(hipMemcpy(ctx->E, h->E, sizeof(double) * N, hipMemcpyHostToDevice));
    if (h->C)
        CK(hipMemcpy(ctx->C, h->C, sizeof(double) * N, hipMemcpyHostToDevice));
    else
        CK(hipMemset(ctx->C, 0, sizeof(double) * N));
    for (size_t c = 0; c < N; c++)
        for (int q = 0; q < nv; q++)
            buf[q * N + c] = h->CS[c * nv + q];
    CK(hipMemcpy(ctx->CS, buf.data(), sizeof(double) * N * nv, hipMemcpyHostToDevice));
    if (h->CMS) {
        for (size_t c = 0; c < N; c++)
            for (int s = 0; s < ns; s++)
                buf[s * N + c] = h->CMS[c * ns + s];
        CK(hipMemcpy(ctx->CMS, buf.data(), sizeof(double) * N * ns, hipMemcpyHostToDevice));
    }
    if (h->Nchem)
        CK(hipMemcpy(ctx->nchem, h->Nchem, sizeof(int) * N, hipMemcpyHostToDevice));
    ctx->has_state = true;
    return EU3D_OK;
}

int eu3d_download_state(eu3d_ctx *ctx, eu3d_state *h)
{
    if (!ctx || !h)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    const size_t N = (size_t)ctx->g.N;
    const int ns = ctx->ns, nv = ctx->nv;
    std::vector<double> prim(N * (7 + ns)), T(N), C(N), cp(N), cs(N * nv), cms(N * ns);
    EU3D_DISPATCH_NS(ns, launch_derive_cp, ctx, ctx->scratch);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(prim.data(), ctx->prim, sizeof(double) * prim.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(T.data(), ctx->T, sizeof(double) * N, hipMemcpyDeviceToHost));
    CK(hipMemcpy(C.data(), ctx->C, sizeof(double) * N, hipMemcpyDeviceToHost));
    CK(hipMemcpy(cp.data(), ctx->scratch, sizeof(double) * N, hipMemcpyDeviceToHost));
    CK(hipMemcpy(cs.data(), ctx->CS, sizeof(double) * cs.size(), hipMemcpyDeviceToHost));
    CK(hipMemcpy(cms.data(), ctx->CMS, sizeof(double) * cms.size(), hipMemcpyDeviceToHost));
    double *dst[7] = {h->P, h->D, h->U, h->V, h->W, h->H, h->Gamma};
    for (int f = 0; f < 7; f++)
        if (dst[f])
            std::memcpy(dst[f], prim.data() + f * N, sizeof(double) * N);
    if (h->T)
        std::memcpy(h->T, T.data(), sizeof(double) * N);
    if (h->E)
        CK(hipMemcpy(h->E, ctx->E, sizeof(double) * N, hipMemcpyDeviceToHost));
    if (h->C)
        std::memcpy(h->C, C.data(), sizeof(double) * N);
    if (h->Cp)
        std::memcpy(h->Cp, cp.data(), sizeof(double) * N);
    const DevMech &m = ctx->hmech;
    for (size_t c = 0; c < N; c++) {
        const double *Y = nullptr;
        double sYM = 0.0;
        for (int s = 0; s < ns; s++) {
            double y = prim[(7 + s) * N + c];
            if (h->Yi)
                h->Yi[c * ns + s] = y;
            sYM += y / m.mw[s];
        }
        (void)Y;
        double Wav = 1.0 / sYM;
        if (h->Wav)
            h->Wav[c] = Wav;
        if (h->Rgas)
            h->Rgas[c] = m.R * 1000 / Wav;
        if (h->Ma) {
            double U = prim[2 * N + c], V = prim[3 * N + c], W = prim[4 * N + c];
            h->Ma[c] = std::sqrt((U * U) + (V * V) + (W * W)) / C[c];
        }
        for (int s = 0; s < ns; s++) {
            double di = cs[s * N + c];
            if (h->Di)
                h->Di[c * ns + s] = di;
            if (h->Mc)
                h->Mc[c * ns + s] = di / m.mw[s] * 1000;
            if (h->CMS)
                h->CMS[c * ns + s] = cms[s * N + c];
        }
        if (h->CS)
            for (int q = 0; q < nv; q++)
                h->CS[c * nv + q] = cs[q * N + c];
    }
    if (h->Nchem)
        CK(hipMemcpy(h->Nchem, ctx->nchem, sizeof(int) * N, hipMemcpyDeviceToHost));
    return EU3D_OK;
}

int eu3d_fill_ghosts(eu3d_ctx *ctx)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    Timer tm(ctx, K_GHOST);
    const Grid &g = ctx->g;
    long long nx = 2LL * g.bc * g.nj * g.nk;
    k_ghost_x<<<blocks_for(nx), kBlock, 0, ctx->stream>>>(g, ctx->nprim, ctx->prim, ctx->T, ctx->E);
    EU3D_DISPATCH_NS(ctx->ns, launch_ghost_y, ctx);
    long long nz = 2LL * g.NI * g.NJ * g.bc;
    k_ghost_z<<<blocks_for(nz), kBlock, 0, ctx->stream>>>(g, ctx->nprim, ctx->prim, ctx->T, ctx->E);
    CK(hipGetLastError());
    return EU3D_OK;
}

int eu3d_cfl_dt(eu3d_ctx *ctx, double cfl, double time, double final_time, double *dt)
{
    if (!ctx || !dt)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    double hx, hy;
    CK(hipMemcpy(&hx, ctx->x, sizeof(double), hipMemcpyDeviceToHost));
    {
        double x1;
        CK(hipMemcpy(&x1, ctx->x + 1, sizeof(double), hipMemcpyDeviceToHost));
        hx = x1 - hx;
    }
    CK(hipMemcpy(&hy, ctx->y, sizeof(double), hipMemcpyDeviceToHost));
    {
        double y1;
        CK(hipMemcpy(&y1, ctx->y + 1, sizeof(double), hipMemcpyDeviceToHost));
        hy = y1 - hy;
    }
    unsigned long long init;
    double big = 1e10;
    std::memcpy(&init, &big, sizeof init);
    unsigned long long *bits = &ctx->cnt->cfl_bits;
    {
        Timer tm(ctx, K_CFL);
        k_set_u64<<<1, 1, 0, ctx->stream>>>(bits, init);
        unsigned int nb = std::min<unsigned int>(blocks_for(ctx->n_int), (unsigned int)ctx->ncu * 8);
        k_cfl<<<nb, kBlock, 0, ctx->stream>>>(ctx->g, ctx->prim, ctx->C, hx, hy, bits);
        CK(hipGetLastError());
    }
    unsigned long long res;
    CK(hipMemcpyAsync(&res, bits, sizeof res, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    double dt_conv;
    std::memcpy(&dt_conv, &res, sizeof dt_conv);
    double d = cfl * dt_conv;
    if (time + d > final_time)
        d = final_time - time;
    *dt = d;
    return EU3D_OK;
}

int eu3d_flux_update(eu3d_ctx *ctx, double dt)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    if (!(dt > 0))
        FAIL(EU3D_ERR_ARG, "dt must be > 0");
    Timer tm(ctx, K_FLUX);
    EU3D_DISPATCH_NS(ctx->ns, launch_flux, ctx, 0, dt);
    CK(hipGetLastError());
    return EU3D_OK;
}

int eu3d_recover_state(eu3d_ctx *ctx)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    Timer tm(ctx, K_RECOVER);
    EU3D_DISPATCH_NS(ctx->ns, launch_recover, ctx);
    CK(hipGetLastError());
    return EU3D_OK;
}

int eu3d_advance_rk3(eu3d_ctx *ctx, double dt)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    if (!(dt > 0))
        FAIL(EU3D_ERR_ARG, "dt must be > 0");
    for (int stage = 1; stage <= 3; stage++) {
        int rc = eu3d_fill_ghosts(ctx);
        if (rc)
            return rc;
        {
            Timer tm(ctx, K_FLUX);
            EU3D_DISPATCH_NS(ctx->ns, launch_flux, ctx, stage, dt);
        }
        rc = eu3d_recover_state(ctx);
        if (rc)
            return rc;
    }
    CK(hipGetLastError());
    return EU3D_OK;
}

int eu3d_chemistry(eu3d_ctx *ctx, double dt, int flags)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    if (!(dt > 0))
        FAIL(EU3D_ERR_ARG, "dt must be > 0");
    if (flags & EU3D_CHEM_SKIP)
        return EU3D_OK;
    Timer tm(ctx, K_CHEM);
    EU3D_DISPATCH_NS(ctx->ns, launch_chem, ctx, dt, flags);
    CK(hipGetLastError());
    return EU3D_OK;
}

int eu3d_step(eu3d_ctx *ctx, int scheme, double dt, int chem_flags)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    int rc;
    if (scheme == EU3D_SCHEME_EE) {
        if ((rc = eu3d_fill_ghosts(ctx)))
            return rc;
        if ((rc = eu3d_flux_update(ctx, dt)))
            return rc;
        if ((rc = eu3d_recover_state(ctx)))
            return rc;
    } else if (scheme == EU3D_SCHEME_TVD_RK3) {
        if ((rc = eu3d_advance_rk3(ctx, dt)))
            return rc;
    } else {
        FAIL(EU3D_ERR_ARG, "unknown time-advance scheme");
    }
    return eu3d_chemistry(ctx, dt, chem_flags);
}

int eu3d_run(eu3d_ctx *ctx, int scheme, double cfl, double dt_fixed, double final_time, int max_steps,
             int chem_flags, double *time, int *steps_done)
{
    if (!ctx || !time)
        return EU3D_ERR_ARG;
    int n = 0;
    const double tol = 1e-12 * std::fabs(final_time);
    int rc = EU3D_OK;
    while (final_time - *time > tol && n < max_steps) {
        double dt;
        if (cfl == 0)
            dt = dt_fixed;
        else if ((rc = eu3d_cfl_dt(ctx, cfl, *time, final_time, &dt)))
            break;
        if (dt > final_time - *time)
            dt = final_time - *time;
        if ((rc = eu3d_step(ctx, scheme, dt, chem_flags)))
            break;
        *time += dt;
        n++;
    }
    if (steps_done)
        *steps_done = n;
    return rc;
}

int eu3d_health(eu3d_ctx *ctx, eu3d_health_info *out)
{
    if (!ctx || !out)
        return EU3D_ERR_ARG;
    ChemCounters c;
    CK(hipMemcpyAsync(&c, ctx->cnt, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    CK(hipStreamSynchronize(ctx->stream));
    out->invalid_nchem = (long long)c.invalid;
    out->nonfinite = (long long)c.nonfinite;
    out->max_nchem = c.max_nchem;
    out->stiff_cells = (long long)c.queue_count;
    out->total_substeps = (long long)c.substeps;
    if (c.invalid || c.nonfinite) {
        ctx->err = "invalid Nchem or non-finite state in the last chemistry call";
        return EU3D_ERR_HEALTH;
    }
    return EU3D_OK;
}

int eu3d_synchronize(eu3d_ctx *ctx)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    CK(hipStreamSynchronize(ctx->stream));
    return EU3D_OK;
}

int eu3d_timing_enable(eu3d_ctx *ctx, int on)
{
    if (!ctx)
        return EU3D_ERR_ARG;
    ctx->timing = on != 0;
    return EU3D_OK;
}

int eu3d_timing_get(eu3d_ctx *ctx, eu3d_timing *o)
{
    if (!ctx || !o)
        return EU3D_ERR_ARG;
    CK(hipStreamSynchronize(ctx->stream));
    resolve_timing(ctx);
    o->flux_ms = ctx->last_ms[K_FLUX];
    o->recover_ms = ctx->last_ms[K_RECOVER];
    o->chem_ms = ctx->last_ms[K_CHEM];
    o->ghost_ms = ctx->last_ms[K_GHOST];
    o->cfl_ms = ctx->last_ms[K_CFL];
    o->flux_total_ms = ctx->total_ms[K_FLUX];
    o->recover_total_ms = ctx->total_ms[K_RECOVER];
    o->chem_total_ms = ctx->total_ms[K_CHEM];
    o->ghost_total_ms = ctx->total_ms[K_GHOST];
    o->cfl_total_ms = ctx->total_ms[K_CFL];
    o->flux_calls = ctx->calls[K_FLUX];
    o->recover_calls = ctx->calls[K_RECOVER];
    o->chem_calls = ctx->calls[K_CHEM];
    o->ghost_calls = ctx->calls[K_GHOST];
    o->cfl_calls = ctx->calls[K_CFL];
    for (int q = 0; q < 5; q++) {
        ctx->total_ms[q] = 0;
        ctx->calls[q] = 0;
    }
    return EU3D_OK;
}

int eu3d_probe_thermo(eu3d_ctx *ctx, const double *T, int n, double R, double *cp, double *h, double *s)
{
    if (!ctx || !T || n < 1 || !cp || !h || !s)
        return EU3D_ERR_ARG;
    int ns = ctx->ns;
    double *d = nullptr;
    CK(hipMalloc(&d, sizeof(double) * (n + 3 * (size_t)n * ns)));
    CK(hipMemcpy(d, T, sizeof(double) * n, hipMemcpyHostToDevice));
    double *dcp = d + n, *dh = dcp + (size_t)n * ns, *ds = dh + (size_t)n * ns;
    EU3D_DISPATCH_NS(ns, launch_probe_thermo, ctx, d, n, R, dcp, dh, ds);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(cp, dcp, sizeof(double) * n * ns, hipMemcpyDeviceToHost));
    CK(hipMemcpy(h, dh, sizeof(double) * n * ns, hipMemcpyDeviceToHost));
    CK(hipMemcpy(s, ds, sizeof(double) * n * ns, hipMemcpyDeviceToHost));
    CK(hipFree(d));
    return EU3D_OK;
}

int eu3d_probe_face_flux(eu3d_ctx *ctx, int direction, double *flux)
{
    if (!ctx || !flux || direction < 1 || direction > 3)
        return EU3D_ERR_ARG;
    if (!ctx->has_state)
        FAIL(EU3D_ERR_STATE, "no state uploaded");
    const size_t N = (size_t)ctx->g.N;
    const int nv = ctx->nv;
    double *d = nullptr;
    CK(hipMalloc(&d, sizeof(double) * N * nv));
    CK(hipMemsetAsync(d, 0, sizeof(double) * N * nv, ctx->stream));
    EU3D_DISPATCH_NS(ctx->ns, launch_probe_flux, ctx, direction - 1, d);
    CK(hipStreamSynchronize(ctx->stream));
    std::vector<double> buf(N * nv);
    CK(hipMemcpy(buf.data(), d, sizeof(double) * N * nv, hipMemcpyDeviceToHost));
    CK(hipFree(d));
    for (size_t c = 0; c < N; c++)
        for (int q = 0; q < nv; q++)
            flux[c * nv + q] = buf[q * N + c];
    return EU3D_OK;
}

int eu3d_probe_rates(eu3d_ctx *ctx, const double *T, const double *Mc, int n, double *kf, double *kb, double *rr,
                     double *wi)
{
    (void)kf;
    (void)kb;
    (void)rr;
    if (!ctx || !T || !Mc || n < 1 || !wi)
        return EU3D_ERR_ARG;
    int ns = ctx->ns;
    double *d = nullptr;
    CK(hipMalloc(&d, sizeof(double) * (n + 2 * (size_t)n * ns)));
    CK(hipMemcpy(d, T, sizeof(double) * n, hipMemcpyHostToDevice));
    CK(hipMemcpy(d + n, Mc, sizeof(double) * n * ns, hipMemcpyHostToDevice));
    EU3D_DISPATCH_NS(ns, launch_probe_rates, ctx, d, d + n, n, d + n + (size_t)n * ns);
    CK(hipStreamSynchronize(ctx->stream));
    CK(hipMemcpy(wi, d + n + (size_t)n * ns, sizeof(double) * n * ns, hipMemcpyDeviceToHost));
    CK(hipFree(d));
    return EU3D_OK;
}

} // extern "C"
