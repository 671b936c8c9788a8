// Stout smearing and its force chain on gfx950 (FP64).
//
//  smearing   U' = exp(-QA) U, QA = (rho/c0) TA(U * C), C = c0 * sum of the 6
//             staples (OpenAcc/stouting.c:27-165, plaquettes.c:196-259,
//             su3_utilities.c:210-239; Morningstar-Peardon hep-lat/0311018)
//  force      Sigma' -> Lambda -> Sigma (stouting.c:171-1302,
//             fermion_force.c:52-162; MP eqs. 57-75)
//
// One thread per link (8 link arrays x sizeh half-sites); the staple loop
// re-reads neighbouring links, which the L2 of the owning XCD serves
// (workgroups are remapped so each XCD walks a contiguous slab).  With about
// 2.5 kflop per link against 192 B of compulsory link traffic these kernels sit
// near the FP64 ridge; everything stays in FP64 as the reference does.
#include "ostp_blas.h"
#include "ostp_su3.h"

#include <cmath>

extern "C" {
// action.h:23 and the staple normalisation of the gauge action (common_defines.h:71-84)
double gl_stout_rho = 0.15;
double gl_topo_rho = 0.0;
}

namespace ostp {

static double g_c_zero = 5.0 * 0.33333333333333333333333;  // tree-level Symanzik C_ZERO
static int g_stout_steps = 2;
static int g_topo_steps = 0;
static int g_topo_action = 0;
static bool g_tlsm = true;

bool gauge_action_tlsm() { return g_tlsm; }

struct LinkGeom {
    int nd[4];
    int sizeh;
};

// the 8 link arrays of a gauge field, contiguous at a fixed stride: a link
// array's address is arithmetic on a per-lane index, not a load from an
// array of pointers
struct GaugePtrs {
    const char* base;  // r0.c0 plane of link array 0; rows at +3k*sizeh, cols at +c*sizeh
    size_t stride;     // bytes between link arrays
    __device__ __forceinline__ const double2* at(int l) const { return (const double2*)(base + (size_t)l * stride); }
};

__device__ __forceinline__ void decode(int idxh, int parity, const LinkGeom& g, int d[4]) {
    const int nd0h = g.nd[0] >> 1;
    int t = idxh / nd0h;
    const int hd0 = idxh - t * nd0h;
    int t2 = t / g.nd[1];
    d[1] = t - t2 * g.nd[1];
    d[3] = t2 / g.nd[2];
    d[2] = t2 - d[3] * g.nd[2];
    d[0] = 2 * hd0 + ((d[1] + d[2] + d[3] + parity) & 1);
}

__device__ __forceinline__ int snum(const int d[4], const LinkGeom& g) {
    return (d[0] + g.nd[0] * (d[1] + g.nd[1] * (d[2] + g.nd[2] * d[3]))) >> 1;
}

// half-site index of x + s*e_mu (+ t*e_nu), periodic
__device__ __forceinline__ int shifted(const int d[4], int mu, int s, int nu, int t, const LinkGeom& g) {
    int e[4] = {d[0], d[1], d[2], d[3]};
    e[mu] += s;
    if (e[mu] < 0) e[mu] += g.nd[mu];
    if (e[mu] >= g.nd[mu]) e[mu] -= g.nd[mu];
    if (t != 0) {
        e[nu] += t;
        if (e[nu] < 0) e[nu] += g.nd[nu];
        if (e[nu] >= g.nd[nu]) e[nu] -= g.nd[nu];
    }
    return snum(e, g);
}

__device__ __forceinline__ void load_link(const double2* __restrict__ U, int sizeh, int i, Su3& m) {
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) m.c[r][c] = U[(3 * r + c) * sizeh + i];
    rebuild3row(m);
}
__device__ __forceinline__ void load_gl3(const double2* __restrict__ U, int sizeh, int i, Su3& m) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) m.c[r][c] = U[(3 * r + c) * sizeh + i];
}
__device__ __forceinline__ void store_rows(double2* __restrict__ U, int sizeh, int i, const Su3& m, int rows) {
    for (int r = 0; r < rows; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) U[(3 * r + c) * sizeh + i] = m.c[r][c];
}

__device__ __constant__ int c_perp[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

// sum of the 6 staples of link (x, mu) times c0 (plaquettes.c:196-259)
__device__ __forceinline__ void staple_sum(const GaugePtrs& G, const LinkGeom& g, const int d[4], int idxh, int mu,
                                           int parity, double c0, Su3& acc) {
    su3_zero(acc);
    const int sh = g.sizeh;
    const int i_pmu = shifted(d, mu, 1, 0, 0, g);
    for (int it = 0; it < 3; it++) {
        const int nu = c_perp[mu][it];
        Su3 A, B, C, T;
        // right: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        load_link(G.at(2 * nu + !parity), sh, i_pmu, A);
        load_link(G.at(2 * mu + !parity), sh, shifted(d, nu, 1, 0, 0, g), B);
        load_link(G.at(2 * nu + parity), sh, idxh, C);
        su3_mul_dag(A, B, T);
        su3_mul_dag(T, C, T);
        rebuild3row(T);
        su3_scale_add(c0, T, acc);
        // left: U_nu(x+mu-nu)^+ U_mu(x-nu)^+ U_nu(x-nu)
        const int i_mnu = shifted(d, nu, -1, 0, 0, g);
        load_link(G.at(2 * nu + parity), sh, shifted(d, mu, 1, nu, -1, g), A);
        load_link(G.at(2 * mu + !parity), sh, i_mnu, B);
        load_link(G.at(2 * nu + !parity), sh, i_mnu, C);
        Su3 AB;
        su3_mul(B, A, AB);      // A^+ B^+ = (B A)^+
        su3_dag_mul(AB, C, T);  // (B A)^+ C
        rebuild3row(T);
        su3_scale_add(c0, T, acc);
    }
}

__device__ __forceinline__ void link_of(int gid, int sizeh, int& dir, int& idxh) {
    dir = gid / sizeh;
    idxh = gid - dir * sizeh;
}
// The link of this thread in the kernels that read neighbour links (staples):
// blocks run over the 8 link arrays fastest, then over tiles of kBlock
// half-sites, and each XCD takes a contiguous range of blocks, so an XCD
// owns a slab of the lattice for every link array and re-reads the slab's
// neighbour links from its own L2 (with the link array slowest, every XCD
// swept the whole lattice for one array and fetched every neighbour link).
// Grid: 8 blocks per tile (nblk_tiles); false past the last half-site.
__device__ __forceinline__ bool link_of_tile(int sizeh, int& dir, int& idxh) {
    const int blk = xcd_block_id();
    dir = blk & 7;
    idxh = (blk >> 3) * kBlock + (int)threadIdx.x;
    return idxh < sizeh;
}


// ============================================================ gauge force
// Tree-level Symanzik (or Wilson) gauge force and action: plaquette staples
// times C_ZERO (plaquettes.c calc_loc_staples_nnptrick_all), rectangle
// staples of types A, B, C times C_ONE (rectangles.c:643-916), with the
// PAR_TEMP defect coefficients K (products of the staple's links, then the
// link's own K, su3_utilities.c:582-604), and ipdot = TA(U * staple)
// (conf_times_staples_ta_part, su3_utilities.c:146-174).

constexpr double kCOne = -0.25 * 0.33333333333333333333333;  // common_defines.h:75

// y = x + a e_mu + b e_nu, periodic (|a|, |b| <= 2 and every extent >= 2)
__device__ __forceinline__ void site_off(const int x[4], int mu, int a, int nu, int b, const LinkGeom& g, int y[4]) {
#pragma unroll
    for (int i = 0; i < 4; i++) y[i] = x[i];
    int v = y[mu] + a;
    v += v < 0 ? g.nd[mu] : 0;
    v -= v >= g.nd[mu] ? g.nd[mu] : 0;
    y[mu] = v;
    v = y[nu] + b;
    v += v < 0 ? g.nd[nu] : 0;
    v -= v >= g.nd[nu] ? g.nd[nu] : 0;
    y[nu] = v;
}

// U_dir(y): rows 0,1 loaded (row 2 is left for the caller to rebuild when a
// product needs it); *k = its K when the layout has one
__device__ __forceinline__ void link2_y(const GaugePtrs& G, const LinkGeom& g, int dir, const int y[4], int pt, Su3& m,
                                        double& k) {
    const int par = (y[0] + y[1] + y[2] + y[3]) & 1;
    const int i = snum(y, g);
    const double2* U = G.at(2 * dir + par);
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) m.c[r][c] = U[(size_t)(3 * r + c) * g.sizeh + i];
    k = pt ? ((const double*)(U + 9 * (size_t)g.sizeh))[i] : 1.0;
}
__device__ __forceinline__ void link_y(const GaugePtrs& G, const LinkGeom& g, int dir, const int y[4], int pt, Su3& m,
                                       double& k) {
    link2_y(G, g, dir, y, pt, m, k);
    rebuild3row(m);
}

// rows 0,1 of A B, A B^+ (A: rows 0,1; B: all rows); O may alias A
__device__ __forceinline__ void mul_r2(const Su3& A, const Su3& B, Su3& O) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
        double2 o[3];
#pragma unroll
        for (int c = 0; c < 3; c++) o[c] = cfma(A.c[r][2], B.c[2][c], cfma(A.c[r][1], B.c[1][c], cmul(A.c[r][0], B.c[0][c])));
#pragma unroll
        for (int c = 0; c < 3; c++) O.c[r][c] = o[c];
    }
}
__device__ __forceinline__ void mul_dag_r2(const Su3& A, const Su3& B, Su3& O) {
#pragma unroll
    for (int r = 0; r < 2; r++) {
        double2 o[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            o[c] = cfma(A.c[r][2], cconj(B.c[c][2]), cfma(A.c[r][1], cconj(B.c[c][1]), cmul(A.c[r][0], cconj(B.c[c][0]))));
#pragma unroll
        for (int c = 0; c < 3; c++) O.c[r][c] = o[c];
    }
}
// rows 0,1 of A^+ B (A, B: all rows); O must not alias A
__device__ __forceinline__ void dag_mul_r2(const Su3& A, const Su3& B, Su3& O) {
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            O.c[r][c] = cfma(cconj(A.c[2][r]), B.c[2][c], cfma(cconj(A.c[1][r]), B.c[1][c], cmul(cconj(A.c[0][r]), B.c[0][c])));
}

// five-link staple term m1^(d1) m2^(d2) m3^+ m4^(d4) m5^(d5) for the three
// shapes of rectangles.c: PPMMM (d = 0,0,1,1,1), PMMMP (0,1,1,1,0), MMMPP (1,1,1,0,0)
enum Shape : int { kPPMMM = 0, kPMMMP = 1, kMMMPP = 2 };
struct Hop {
    int dir, a, b;  // link U_dir at x + a e_mu + b e_nu
};
// rows 0,1 of m^+ (m: all rows)
__device__ __forceinline__ void dag_r2(const Su3& m, Su3& O) {
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) O.c[r][c] = cconj(m.c[c][r]);
}
// acc += s T (transposed = 0) or s T^+ (transposed = 1), T with all rows
__device__ __forceinline__ void add_term(double s, const Su3& T, int transposed, Su3& acc) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double2 v = transposed ? cconj(T.c[c][r]) : T.c[r][c];
            acc.c[r][c] = make_double2(acc.c[r][c].x + s * v.x, acc.c[r][c].y + s * v.y);
        }
}
// The links of a wave's staple terms.  A wave of the gauge force covers 64
// sites of one link array (direction mu, parity p) and one plane nu, so mu,
// nu, p, the strides and the link array of every hop are wave-uniform and
// live in scalar registers; a lane keeps its site's lexicographic index and
// its mu, nu coordinates, and a neighbour x + a e_mu + b e_nu is
// (lex + a s_mu + b s_nu, wrapped) / 2 in the link array of parity
// p + a + b.  Loads are a scalar base plus a 32-bit lane offset.
struct PlaneCtx {
    const char* base;
    size_t stride;
    unsigned plane;  // bytes per colour plane (16 sizeh)
    int mu, nu, p, sm, sn, ndm, ndn;
    int lex, xm, xn;
};
__device__ __forceinline__ int pc_idx(const PlaneCtx& c, int a, int b) {
    const int ym = c.xm + a, yn = c.xn + b;
    int l = c.lex + a * c.sm + b * c.sn;
    l += ym < 0 ? c.ndm * c.sm : (ym >= c.ndm ? -c.ndm * c.sm : 0);
    l += yn < 0 ? c.ndn * c.sn : (yn >= c.ndn ? -c.ndn * c.sn : 0);
    return l >> 1;
}
// rows 0,1 of U_d(x + a e_mu + b e_nu), d = mu (which = 0) or nu (1); *k its K
__device__ __forceinline__ void pc_link2(const PlaneCtx& c, int which, int a, int b, int pt, Su3& m, double& k) {
    const int arr = __builtin_amdgcn_readfirstlane(2 * (which ? c.nu : c.mu) + ((c.p + a + b) & 1));
    const char* U = c.base + (size_t)arr * c.stride;
    const unsigned i16 = (unsigned)pc_idx(c, a, b) * 16u;
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int col = 0; col < 3; col++)
            m.c[r][col] = *(const double2*)(U + (i16 + (unsigned)(3 * r + col) * c.plane));
    k = pt ? *(const double*)(U + (size_t)9 * c.plane + (i16 >> 1)) : 1.0;
}
__device__ __forceinline__ void pc_link(const PlaneCtx& c, int which, int a, int b, int pt, Su3& m, double& k) {
    pc_link2(c, which, a, b, pt, m, k);
    rebuild3row(m);
}

// The plane kernel's link source: rows 0,1 and K of the links of a tile of
// the plane (mu, nu) with a two-site halo, in LDS.  Region site (i, j) is
// tile site (i - 2, j - 2) (i along mu, j along nu); slot d ns / 2 + j rw + i
// holds its link along mu (d = 0) or nu (d = 1), so that lanes on
// consecutive sites read consecutive 16-B entries.  A staple term of a link
// along mu takes offsets a (along mu) and b (along nu) as in PlaneCtx; for a
// link along nu (own = 1) the roles of the two directions swap.
struct LdsTile {
    const double2* e;  // 6 arrays of ns entries: row r, column c at (3 r + c) ns
    const double* k;   // ns entries
    int ns, rw;
    int i0, j0;        // region coordinates of the link's site
    int own;
};
__device__ __forceinline__ int lt_slot(const LdsTile& c, int which, int a, int b) {
    const int di = c.own ? b : a, dj = c.own ? a : b;
    return (c.own ? 1 - which : which) * (c.ns >> 1) + (c.j0 + dj) * c.rw + (c.i0 + di);
}

// rows 0,1 (src_link2) or all rows (src_link) of a staple term's link and its K
__device__ __forceinline__ void src_link2(const PlaneCtx& c, int which, int a, int b, int pt, Su3& m, double& k) {
    pc_link2(c, which, a, b, pt, m, k);
}
__device__ __forceinline__ void src_link(const PlaneCtx& c, int which, int a, int b, int pt, Su3& m, double& k) {
    pc_link(c, which, a, b, pt, m, k);
}
__device__ __forceinline__ void src_link2(const LdsTile& c, int which, int a, int b, int pt, Su3& m, double& k) {
    const int s = lt_slot(c, which, a, b);
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int col = 0; col < 3; col++) m.c[r][col] = c.e[(3 * r + col) * c.ns + s];
    k = pt ? c.k[s] : 1.0;
}
__device__ __forceinline__ void src_link(const LdsTile& c, int which, int a, int b, int pt, Su3& m, double& k) {
    src_link2(c, which, a, b, pt, m, k);
    rebuild3row(m);
}

// Each term is a chain of SU(3) products that carries only rows 0,1 and
// rebuilds row 2 at the end.  MMMPP is formed as its conjugate
// m5^+ m4^+ m3 m2 m1, a chain of the same kind, and added transposed.
template <int SH>
__device__ __forceinline__ void five(const PlaneCtx& pc, const Hop* h, int pt, double c1, Su3& acc) {
    double k = 1.0;
    auto ld = [&](int j, Su3& m, bool full) {
        double kj;
        pc_link2(pc, h[j].dir, h[j].a, h[j].b, pt, m, kj);
        if (full) rebuild3row(m);
        k *= kj;
    };
    // dagger pattern of the chain, in chain order
    constexpr int d1 = SH == kPPMMM ? 0 : SH == kPMMMP ? 0 : 1;
    constexpr int d2 = SH == kPPMMM ? 0 : 1;
    constexpr int d4 = SH == kPMMMP ? 1 : SH == kPPMMM ? 1 : 0;
    constexpr int d5 = SH == kPMMMP ? 0 : SH == kPPMMM ? 1 : 0;
    Su3 T, M;
    if (SH != kMMMPP) {
        ld(0, T, false);
        ld(1, M, true);
        if (d2) mul_dag_r2(T, M, T); else mul_r2(T, M, T);
        ld(2, M, true);
        mul_dag_r2(T, M, T);
        ld(3, M, true);
        if (d4) mul_dag_r2(T, M, T); else mul_r2(T, M, T);
        ld(4, M, true);
        if (d5) mul_dag_r2(T, M, T); else mul_r2(T, M, T);
    } else {
        (void)d1;
        ld(4, M, true);
        dag_r2(M, T);
        ld(3, M, true);
        mul_dag_r2(T, M, T);
        ld(2, M, true);
        mul_r2(T, M, T);
        ld(1, M, true);
        mul_r2(T, M, T);
        ld(0, M, true);
        mul_r2(T, M, T);
    }
    rebuild3row(T);
    add_term(k * c1, T, SH == kMMMPP, acc);
}

// rectangle staple terms A right/left, B right/left, C right/left
// (rectangles.c:691-718, 780-807, 865-892): the five links and the shape
__device__ __constant__ Hop c_rect_hops[6][5] = {
    {{1, 1, 0}, {1, 1, 1}, {0, 0, 2}, {1, 0, 1}, {1, 0, 0}},
    {{1, 1, -1}, {1, 1, -2}, {0, 0, -2}, {1, 0, -2}, {1, 0, -1}},
    {{0, 1, 0}, {1, 2, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}},
    {{0, 1, 0}, {1, 2, -1}, {0, 1, -1}, {0, 0, -1}, {1, 0, -1}},
    {{1, 1, 0}, {0, 0, 1}, {0, -1, 1}, {1, -1, 0}, {0, -1, 0}},
    {{1, 1, -1}, {0, 0, -1}, {0, -1, -1}, {1, -1, -1}, {0, -1, 0}},
};
__device__ __constant__ int c_rect_shape[6] = {kPPMMM, kMMMPP, kPPMMM, kPMMMP, kPMMMP, kMMMPP};

// the staple terms of link (x, mu) in the plane (mu, nu), in the reference's
// order within the plane: plaquettes right and left times C_ZERO (plaquettes.c:
// 83-195), rectangles of types A, B, C right and left times C_ONE (rectangles.c:
// 691-718, 780-807, 865-892); K products of each term's links when pt
// A link of a staple term: U_d(x + a e_mu + b e_nu), d = mu (which = 0) or
// nu (1), daggered when dag.
struct HopD {
    int which, a, b, dag;
};
// T = L0 L1 ... L(N-1) with all rows; *k the product of the links' K
template <int N, class Src>
__device__ __forceinline__ void chain(const Src& pc, const HopD (&h)[N], int pt, Su3& T, double& k) {
    Su3 M;
    double kj;
    if (h[0].dag) {
        src_link(pc, h[0].which, h[0].a, h[0].b, pt, M, kj);
        dag_r2(M, T);
    } else {
        src_link2(pc, h[0].which, h[0].a, h[0].b, pt, T, kj);
    }
    k = kj;
#pragma unroll
    for (int j = 1; j < N; j++) {
        src_link(pc, h[j].which, h[j].a, h[j].b, pt, M, kj);
        if (h[j].dag) mul_dag_r2(T, M, T); else mul_r2(T, M, T);
        k *= kj;
    }
    rebuild3row(T);
}
// sl[e kGfLinks] (e = 2 (3 r + c) + re/im) += s T: a thread's staple sum in its
// own LDS slots, which keeps 36 registers free for the product chains
constexpr int kGfLinks = 64;
__device__ __forceinline__ void lds_axpy(double s, const Su3& T, double* sl) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const int e = 2 * (3 * r + c);
            sl[e * kGfLinks] = fma(s, T.c[r][c].x, sl[e * kGfLinks]);
            sl[(e + 1) * kGfLinks] = fma(s, T.c[r][c].y, sl[(e + 1) * kGfLinks]);
        }
}
// acc += s T (all rows)
__device__ __forceinline__ void axpy3(double s, const Su3& T, Su3& acc) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++)
            acc.c[r][c] = make_double2(fma(s, T.c[r][c].x, acc.c[r][c].x), fma(s, T.c[r][c].y, acc.c[r][c].y));
}

// The staple terms of link (x, mu) on one side of the plane (mu, nu): the
// plaquette a m n and the rectangles of types A, B, C on that side
// (plaquettes.c:83-195, rectangles.c:691-718, 780-807, 865-892), each
// times C_ZERO or C_ONE and the K of its links.  On the upper side
//   a = U_nu(x+mu), m = U_mu(x+nu)^+, n = U_nu(x)^+,
//   A = a Q0 n, B = Q2 m n, C = a m Q4 with
//   Q0 = U_nu(x+mu+nu) U_mu(x+2nu)^+ U_nu(x+nu)^+,
//   Q2 = U_mu(x+mu) U_nu(x+2mu) U_mu(x+mu+nu)^+,
//   Q4 = U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu),
// and the lower side is its mirror in nu (a = U_nu(x+mu-nu)^+,
// m = U_mu(x-nu)^+, n = U_nu(x-nu)).  The four terms share factors:
//   (c0 a m + c1 a Q0 + c1 Q2 m) n + c1 (a m) Q4,
// 14 link loads and 207 complex multiply-adds instead of 18 and 252 for
// the terms taken one by one; the sum is the same up to rounding.
// c_side[side] = {a, m, n, Q0[3], Q2[3], Q4[3]}
__device__ __constant__ HopD c_side[2][12] = {
    {{1, 1, 0, 0}, {0, 0, 1, 1}, {1, 0, 0, 1},
     {1, 1, 1, 0}, {0, 0, 2, 1}, {1, 0, 1, 1},
     {0, 1, 0, 0}, {1, 2, 0, 0}, {0, 1, 1, 1},
     {0, -1, 1, 1}, {1, -1, 0, 1}, {0, -1, 0, 0}},
    {{1, 1, -1, 1}, {0, 0, -1, 1}, {1, 0, -1, 0},
     {1, 1, -2, 1}, {0, 0, -2, 1}, {1, 0, -2, 0},
     {0, 1, 0, 0}, {1, 2, -1, 1}, {0, 1, -1, 1},
     {0, -1, -1, 1}, {1, -1, -1, 0}, {0, -1, 0, 0}},
};

// one side (c_side[side]) of the plane's staple terms; add(s, T) receives
// each product to be summed with its coefficient
template <class Src, class Add>
__device__ __forceinline__ void plane_side(const Src& pc, const HopD* d, int pt, int tlsm, double c0, Add add) {
    Su3 G, T;
    double k;
    su3_zero(G);
    {
        // a m: into G with c0, and with Q4 into S
        Su3 am;
        double k_am;
        const HopD h[2] = {d[0], d[1]};
        chain(pc, h, pt, am, k_am);
        axpy3(c0 * k_am, am, G);
        if (tlsm) {
            const HopD h4[3] = {d[9], d[10], d[11]};
            chain(pc, h4, pt, T, k);
            Su3 X;
            mul_r2(am, T, X);
            rebuild3row(X);
            add(kCOne * k_am * k, X);
        }
    }
    if (tlsm) {
        const HopD h0[4] = {d[0], d[3], d[4], d[5]};
        chain(pc, h0, pt, T, k);
        axpy3(kCOne * k, T, G);
        const HopD h2[4] = {d[6], d[7], d[8], d[1]};
        chain(pc, h2, pt, T, k);
        axpy3(kCOne * k, T, G);
    }
    // (...) n
    Su3 n;
    double kn;
    src_link(pc, d[2].which, d[2].a, d[2].b, pt, T, kn);
    if (d[2].dag) {
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) n.c[r][c] = cconj(T.c[c][r]);
    } else {
        n = T;
    }
    su3_mul(G, n, T);
    add(kn, T);
}

__device__ __forceinline__ void plane_staple(const PlaneCtx& pc, int pt, int tlsm, double c0, double* sl) {
#pragma unroll 1
    for (int side = 0; side < 2; side++)
        plane_side(pc, c_side[side], pt, tlsm, c0, [&](double s, const Su3& T) { lds_axpy(s, T, sl); });
}

// Gauge force, one thread per (link, plane): a block is three waves over the
// same 64 links, wave w summing the staple terms of plane nu = perp(mu)[w].
// Each wave forms TA(U S_w) (TA is linear, so the planes' parts add up to
// TA(U S)); the three parts (and, when requested, the staples) are summed
// through LDS in plane order and multiplied by K_mu(x).  Wave 0 then writes
// ipdot and, with mom, the momentum step of mom_sum_mult
// (su3_utilities.h:1383-1394).
__global__ __launch_bounds__(3 * kGfLinks, 2) void k_gauge_force(GaugePtrs G, double2* __restrict__ stap,
                                                             size_t su3_stride, void* __restrict__ ipdot, LinkGeom g,
                                                             int pt, int tlsm, double c0, void* __restrict__ mom,
                                                             double fact) {
    __shared__ double red[3][18][kGfLinks];
    const int lane = (int)threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
    // A block covers 64 half-sites of one link array (dir uniform).  Logical
    // blocks run over the link array fastest, then t, then the 64-site tile
    // of the t-slice, and each XCD takes a contiguous range of them: an XCD
    // owns a slab of tiles for every t and every link array, and the blocks
    // it runs together cover one tile at neighbouring t, so the neighbour
    // links of a staple are re-read from its L2 (with each XCD sweeping the
    // whole lattice for one link array, every XCD fetched every neighbour
    // link from memory: 22 GB per launch at 48^3x12).  A t-slice that is not
    // a whole number of tiles keeps the plain half-site order.
    const int blk = xcd_block_id();
    const int dir = __builtin_amdgcn_readfirstlane(blk & 7);
    const int rest = blk >> 3;
    const int s3 = g.sizeh / g.nd[3];
    int idxh;
    if (s3 % kGfLinks == 0) {
        const int tile = rest / g.nd[3], tt = rest - tile * g.nd[3];
        idxh = tt * s3 + tile * kGfLinks + lane;
    } else {
        idxh = rest * kGfLinks + lane;
    }
    const bool live = idxh < g.sizeh;
    const int mu = dir >> 1;
    double* sl = &red[w][0][lane];
#pragma unroll
    for (int e = 0; e < 18; e++) sl[e * kGfLinks] = 0.0;
    Tamat q = {};
    if (live) {
        int x[4];
        decode(idxh, dir & 1, g, x);
        PlaneCtx pc;
        pc.base = G.base;
        pc.stride = G.stride;
        pc.plane = 16u * (unsigned)g.sizeh;
        pc.mu = mu;
        pc.nu = __builtin_amdgcn_readfirstlane(c_perp[mu][w]);
        pc.p = dir & 1;
        const int st[4] = {1, g.nd[0], g.nd[0] * g.nd[1], g.nd[0] * g.nd[1] * g.nd[2]};
        auto sel = [](const int* v, int i) { return i == 0 ? v[0] : i == 1 ? v[1] : i == 2 ? v[2] : v[3]; };
        pc.sm = sel(st, pc.mu);
        pc.sn = sel(st, pc.nu);
        pc.ndm = sel(g.nd, pc.mu);
        pc.ndn = sel(g.nd, pc.nu);
        pc.lex = x[0] + st[1] * x[1] + st[2] * x[2] + st[3] * x[3];
        pc.xm = sel(x, pc.mu);
        pc.xn = sel(x, pc.nu);
        plane_staple(pc, pt, tlsm, c0, sl);
        Su3 S, U, X;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++)
                S.c[r][c] = make_double2(sl[2 * (3 * r + c) * kGfLinks], sl[(2 * (3 * r + c) + 1) * kGfLinks]);
        double k;
        pc_link(pc, 0, 0, 0, pt, U, k);
        su3_mul(U, S, X);
        ta_part(X, 1.0, q);
    }
    const double kmu = (live && pt) ? ((const double*)(G.at(dir) + 9 * (size_t)g.sizeh))[idxh] : 1.0;
    if (stap) {
        // the three planes' staple sums are in red
        __syncthreads();
        if (w == 0 && live) {
            double2* O = (double2*)((char*)stap + dir * su3_stride);
#pragma unroll
            for (int e = 0; e < 9; e++) {
                const double re = kmu * (red[0][2 * e][lane] + red[1][2 * e][lane] + red[2][2 * e][lane]);
                const double im = kmu * (red[0][2 * e + 1][lane] + red[1][2 * e + 1][lane] + red[2][2 * e + 1][lane]);
                O[(size_t)e * g.sizeh + idxh] = make_double2(re, im);
            }
        }
    }
    __syncthreads();
    red[w][0][lane] = q.c01.x;
    red[w][1][lane] = q.c01.y;
    red[w][2][lane] = q.c02.x;
    red[w][3][lane] = q.c02.y;
    red[w][4][lane] = q.c12.x;
    red[w][5][lane] = q.c12.y;
    red[w][6][lane] = q.ic00;
    red[w][7][lane] = q.ic11;
    __syncthreads();
    if (w == 0 && live) {
        double t[8];
#pragma unroll
        for (int j = 0; j < 8; j++) t[j] = kmu * (red[0][j][lane] + red[1][j][lane] + red[2][j][lane]);
        Tamat o;
        o.c01 = make_double2(t[0], t[1]);
        o.c02 = make_double2(t[2], t[3]);
        o.c12 = make_double2(t[4], t[5]);
        o.ic00 = t[6];
        o.ic11 = t[7];
        ta_store(ipdot, dir, g.sizeh, idxh, o);
        if (mom) mom_step(mom, dir, g.sizeh, idxh, o, fact);
    }
}

// Gauge force by planes: one launch per plane (mu, nu), mu < nu, in the order
// (0,1) (0,2) (0,3) (1,2) (1,3) (2,3).  A block takes a TM x TN tile of the
// plane at fixed coordinates along the other two directions, loads rows 0,1
// and K of the tile's links along mu and nu with a two-site halo into LDS
// (every staple term of the plane stays within it), and forms the plane's
// terms for the tile's links along mu and along nu, one side of the plane
// per thread, with every link operand read from LDS (the per-link kernel
// above fetches 84 links per link from L2).  TA(U S) is linear, so a plane
// adds K(x, d) TA(U S_plane) to ipdot (the first plane of a direction
// stores it) and the last plane of a direction applies the momentum step
// with the sum.  Blocks run over the tiles of one (other-coordinate) slice
// fastest, so that an XCD's consecutive tiles share their halos in its L2.
__device__ __forceinline__ int sel4(int k, int v0, int v1, int v2, int v3) {
    return k == 0 ? v0 : k == 1 ? v1 : k == 2 ? v2 : v3;
}
template <int TM, int TN>
__global__ __launch_bounds__(4 * TM * TN, 2) void k_gf_plane(GaugePtrs G, double2* __restrict__ stap,
                                                             size_t su3_stride, void* __restrict__ ipdot,
                                                             void* __restrict__ mom, LinkGeom g, int mu, int nu,
                                                             int pt, int tlsm, double c0, double fact, int flags) {
    constexpr int RW = TM + 4, RH = TN + 4, NS = 2 * RW * RH, NT = TM * TN;
    __shared__ double2 le[6][NS];
    __shared__ double lk[NS];
    __shared__ double accl[18][4 * NT];  // each thread's side sum (keeps 36 registers free)
    const int oa = (mu != 0 && nu != 0) ? 0 : ((mu != 1 && nu != 1) ? 1 : 2);  // the other two directions
    const int nd0 = g.nd[0], nd1 = g.nd[1], nd2 = g.nd[2], nd3 = g.nd[3];
    const int ndm = sel4(mu, nd0, nd1, nd2, nd3), ndn = sel4(nu, nd0, nd1, nd2, nd3);
    const int nda = sel4(oa, nd0, nd1, nd2, nd3);
    // block -> tile.  Consecutive blocks (the same XCD) take neighbouring
    // tiles: along mu, then nu, then the slice when the plane holds direction
    // 0; otherwise direction 0 (which the half-site index runs along, so a
    // 128-B line holds 16 consecutive x of one parity) fastest, so that the
    // blocks sharing the lines of a tile's links and ipdot run together and
    // the lines are fetched once into the XCD's L2.
    int r = xcd_block_id();
    const int tiles_m = ndm / TM, tiles_n = ndn / TN;
    int tm, tn, xa, xb;
    if (mu == 0) {
        tm = r % tiles_m;
        r /= tiles_m;
        tn = r % tiles_n;
        r /= tiles_n;
        xa = r % nda;
        xb = r / nda;
    } else {
        xa = r % nda;  // oa == 0
        r /= nda;
        tm = r % tiles_m;
        r /= tiles_m;
        tn = r % tiles_n;
        xb = r / tiles_n;
    }
    int o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = k == mu ? tm * TM : k == nu ? tn * TN : k == oa ? xa : xb;
    // links of the region into LDS
    for (int s = threadIdx.x; s < NS; s += 4 * NT) {
        const int d = s >= NS / 2, site = s - d * (NS / 2), i = site % RW, j = site / RW;
        int y[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int n = sel4(k, nd0, nd1, nd2, nd3);
            int v = o[k] + (k == mu ? i - 2 : 0) + (k == nu ? j - 2 : 0);
            v += v < 0 ? n : 0;
            v -= v >= n ? n : 0;
            y[k] = v;
        }
        const int lex = y[0] + nd0 * (y[1] + nd1 * (y[2] + nd2 * y[3]));
        const int arr = 2 * (d ? nu : mu) + ((y[0] + y[1] + y[2] + y[3]) & 1);
        const double2* U = G.at(arr);
        const size_t h = (size_t)(lex >> 1);
#pragma unroll
        for (int q = 0; q < 6; q++) le[q][s] = U[(size_t)q * g.sizeh + h];
        lk[s] = pt ? ((const double*)(U + (size_t)9 * g.sizeh))[h] : 1.0;
    }
    __syncthreads();
    // thread t = (2 dsel + side) NT + item: the link along mu (dsel 0) or nu
    // (dsel 1) of tile site item, the terms on the upper (side 0) or lower
    // (side 1) side of the plane
    const int t = (int)threadIdx.x;
    const int item = t % NT, w = t / NT;
    const int dsel = w >> 1, side = w & 1;
    const int im = item % TM, jn = item / TM;
    LdsTile src{&le[0][0], lk, NS, RW, im + 2, jn + 2, dsel};
#pragma unroll
    for (int e = 0; e < 18; e++) accl[e][t] = 0.0;
    plane_side(src, c_side[side], pt, tlsm, c0, [&](double sc, const Su3& T) {
#pragma unroll
        for (int rr = 0; rr < 3; rr++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const int e = 2 * (3 * rr + cc);
                accl[e][t] = fma(sc, T.c[rr][cc].x, accl[e][t]);
                accl[e + 1][t] = fma(sc, T.c[rr][cc].y, accl[e + 1][t]);
            }
    });
    __syncthreads();
    if (side != 0) return;
    // the plane's staple of the link: both sides' sums (the lower side's
    // thread is t + NT)
    Su3 S, U, X;
#pragma unroll
    for (int rr = 0; rr < 3; rr++)
#pragma unroll
        for (int cc = 0; cc < 3; cc++) {
            const int e = 2 * (3 * rr + cc);
            S.c[rr][cc] = make_double2(accl[e][t] + accl[e][t + NT], accl[e + 1][t] + accl[e + 1][t + NT]);
        }
    double ku;
    src_link(src, 0, 0, 0, pt, U, ku);
    su3_mul(U, S, X);
    Tamat q = {};
    ta_part(X, ku, q);
    int y[4];
#pragma unroll
    for (int k = 0; k < 4; k++) y[k] = o[k] + (k == mu ? im : 0) + (k == nu ? jn : 0);
    const int lex = y[0] + nd0 * (y[1] + nd1 * (y[2] + nd2 * y[3]));
    const int arr = 2 * (dsel ? nu : mu) + ((y[0] + y[1] + y[2] + y[3]) & 1);
    const int idxh = lex >> 1;
    const int first = flags & (dsel ? 4 : 1), last = flags & (dsel ? 8 : 2);
    if (!first) {
        Tamat prev;
        ta_load(ipdot, arr, g.sizeh, idxh, prev);
        q.c01 = cadd(q.c01, prev.c01);
        q.c02 = cadd(q.c02, prev.c02);
        q.c12 = cadd(q.c12, prev.c12);
        q.ic00 += prev.ic00;
        q.ic11 += prev.ic11;
    }
    ta_store(ipdot, arr, g.sizeh, idxh, q);
    if (last && mom) mom_step(mom, arr, g.sizeh, idxh, q, fact);
    if (stap) {
        // staples times K(x, d), summed over the planes (calc_ipdot_gauge_soloopenacc's output)
        double2* O = (double2*)((char*)stap + (size_t)arr * su3_stride) + idxh;
#pragma unroll
        for (int rr = 0; rr < 3; rr++)
#pragma unroll
            for (int cc = 0; cc < 3; cc++) {
                const size_t e = (size_t)(3 * rr + cc) * g.sizeh;
                double2 v = make_double2(ku * S.c[rr][cc].x, ku * S.c[rr][cc].y);
                if (!first) v = cadd(v, O[e]);
                O[e] = v;
            }
    }
}

// per site (both parities): sum over the six planes of Re tr of the
// plaquette (kind 0, plaquettes.c:14-78 times K) or of the 2x1 + 1x2
// rectangles (kind 1, rectangles.c:16-160 times K); block partials
__global__ __launch_bounds__(kBlock) void k_action_sum(GaugePtrs G, LinkGeom g, int pt, int kind, SolverCtl ctl) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    double v = 0.0;
    if (gid < 2 * g.sizeh) {
        const int parity = gid / g.sizeh, idxh = gid - parity * g.sizeh;
        int x[4];
        decode(idxh, parity, g, x);
#pragma unroll 1
        for (int mu = 0; mu < 3; mu++)
#pragma unroll 1
            for (int nu = mu + 1; nu < 4; nu++) {
                int y[4];
                Su3 A, B, C, T;
                double ka, kb, kc, kd;
                if (kind == 0) {
                    // A = U_mu(x), B = U_nu(x+mu), C = U_mu(x+nu), D = U_nu(x): Re tr(A B C^+ D^+)
                    link_y(G, g, mu, x, pt, A, ka);
                    site_off(x, mu, 1, nu, 0, g, y);
                    link_y(G, g, nu, y, pt, B, kb);
                    site_off(x, mu, 0, nu, 1, g, y);
                    link_y(G, g, mu, y, pt, C, kc);
                    Su3 D;
                    link_y(G, g, nu, x, pt, D, kd);
                    su3_mul(A, B, T);
                    su3_mul_dag(T, C, T);
                    su3_mul_dag(T, D, T);
                    v += (T.c[0][0].x + T.c[1][1].x + T.c[2][2].x) * (ka * kb * kc * kd);
                } else {
                    const Hop r21[6] = {{0, 0, 0}, {0, 1, 0}, {1, 2, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}};
                    const Hop r12[6] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 2}, {1, 0, 1}, {1, 0, 0}};
#pragma unroll
                    for (int shape = 0; shape < 2; shape++) {
                        const Hop* h = shape == 0 ? r21 : r12;
                        double k = 1.0;
#pragma unroll
                        for (int j = 0; j < 6; j++) {
                            Su3 M;
                            site_off(x, mu, h[j].a, nu, h[j].b, g, y);
                            double kj;
                            link_y(G, g, h[j].dir == 0 ? mu : nu, y, pt, j == 0 ? T : M, kj);
                            k *= kj;
                            if (j == 1 || j == 2) su3_mul(T, M, T);
                            else if (j >= 3) su3_mul_dag(T, M, T);
                        }
                        v += (T.c[0][0].x + T.c[1][1].x + T.c[2][2].x) * k;
                    }
                }
            }
    }
    block_partial<1>(v, 0.0, ctl);
}

// Fused smearing: U'(rows 0,1) = exp(-QA) U; optional side outputs of the
// reference's parking arrays (staples, QA, exp rows 0,1).
__global__ __launch_bounds__(kBlock) void k_stout(GaugePtrs G, double2* __restrict__ uout, size_t su3_stride,
                                                 LinkGeom g, double rho, double c0, double2* __restrict__ stap_out,
                                                 void* __restrict__ qa_out, double2* __restrict__ exp_out) {
    int dir, idxh;
    if (!link_of_tile(g.sizeh, dir, idxh)) return;
    const int mu = dir >> 1, parity = dir & 1;
    int d[4];
    decode(idxh, parity, g, d);
    Su3 S, U, X, E, O;
    staple_sum(G, g, d, idxh, mu, parity, c0, S);
    load_link(G.at(dir), g.sizeh, idxh, U);
    su3_mul(U, S, X);
    Tamat q;
    ta_part(X, rho / c0, q);
    ch_exp_minus_qa(q, E);
    // rows 0,1 of E * U
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmul(E.c[r][0], U.c[0][c]);
            s = cfma(E.c[r][1], U.c[1][c], s);
            s = cfma(E.c[r][2], U.c[2][c], s);
            O.c[r][c] = s;
        }
    double2* Uo = (double2*)((char*)uout + dir * su3_stride);
    store_rows(Uo, g.sizeh, idxh, O, 2);
    if (stap_out) store_rows((double2*)((char*)stap_out + dir * su3_stride), g.sizeh, idxh, S, 3);
    if (qa_out) ta_store(qa_out, dir, g.sizeh, idxh, q);
    if (exp_out) store_rows((double2*)((char*)exp_out + dir * su3_stride), g.sizeh, idxh, E, 2);
}

// staples accumulated into `stap` (calc_loc_staples_nnptrick_all_onlyferms adds)
__global__ __launch_bounds__(kBlock) void k_staples(GaugePtrs G, double2* __restrict__ stap, size_t su3_stride,
                                                   LinkGeom g, double c0) {
    int dir, idxh;
    if (!link_of_tile(g.sizeh, dir, idxh)) return;
    int d[4];
    decode(idxh, dir & 1, g, d);
    Su3 S, old;
    staple_sum(G, g, d, idxh, dir >> 1, dir & 1, c0, S);
    double2* P = (double2*)((char*)stap + dir * su3_stride);
    load_gl3(P, g.sizeh, idxh, old);
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) old.c[r][c] = cadd(old.c[r][c], S.c[r][c]);
    store_rows(P, g.sizeh, idxh, old, 3);
}

// QA = (rho/c0) TA(U * stap) (RHO_times_conf_times_staples_ta_part)
__global__ __launch_bounds__(kBlock) void k_rho_ta(GaugePtrs G, const double2* __restrict__ stap, size_t su3_stride,
                                                  void* __restrict__ qa, LinkGeom g, double rho, double c0) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * g.sizeh) return;
    int dir, idxh;
    link_of(gid, g.sizeh, dir, idxh);
    Su3 U, S, X;
    load_link(G.at(dir), g.sizeh, idxh, U);
    load_gl3((const double2*)((const char*)stap + dir * su3_stride), g.sizeh, idxh, S);
    su3_mul(U, S, X);
    Tamat q;
    ta_part(X, rho / c0, q);
    ta_store(qa, dir, g.sizeh, idxh, q);
}

// U_out(rows 0,1) = exp(-QA) U (exp_minus_QA_times_conf); exp rows 0,1 to exp_aux
__global__ __launch_bounds__(kBlock) void k_exp_times_conf(GaugePtrs G, const void* __restrict__ qa,
                                                          double2* __restrict__ uout, double2* __restrict__ exp_aux,
                                                          size_t su3_stride, LinkGeom g) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * g.sizeh) return;
    int dir, idxh;
    link_of(gid, g.sizeh, dir, idxh);
    Tamat q;
    ta_load(qa, dir, g.sizeh, idxh, q);
    Su3 E, U, O;
    ch_exp_minus_qa(q, E);
    load_link(G.at(dir), g.sizeh, idxh, U);
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double2 s = cmul(E.c[r][0], U.c[0][c]);
            s = cfma(E.c[r][1], U.c[1][c], s);
            s = cfma(E.c[r][2], U.c[2][c], s);
            O.c[r][c] = s;
        }
    store_rows((double2*)((char*)uout + dir * su3_stride), g.sizeh, idxh, O, 2);
    if (exp_aux) store_rows((double2*)((char*)exp_aux + dir * su3_stride), g.sizeh, idxh, E, 2);
}

// ------------------------------------------------------------- force chain
// Lambda of one link (compute_loc_Lambda, stouting.c:171-512): the derivative
// coefficients b_{1j}, b_{2j} of Morningstar-Peardon eqs. 57-58 (small-c1
// series as in NISSA), B_k = b_k0 + b_k1 Q + b_k2 Q^2, tr_k = Tr(Sigma' B_k U),
// Gamma = tr1 Q + tr2 Q^2 + f1 U Sigma' + f2 (Q U Sigma' + U Sigma' Q),
// Lambda = (Gamma + Gamma^+)/2 - Tr(Gamma + Gamma^+)/6.
__device__ void lambda_local(const Tamat& q, const Su3& SP, const Su3& U, Thmat& L) {
    double c0 = det_i_times_qa(q);
    const double c1 = 0.5 * tr_i_times_qa_sq(q);
    const double c0max = 2 * pow(c1 / 3, 1.5);
    double2 f0, f1, f2, b10, b11, b12, b20, b21, b22;
    if (c1 < 4e-3) {
        ch_coefficients(q, f0, f1, f2);
        b10 = make_double2(0.0, c0 / 120 * (1 - c1 / 21));
        b11 = make_double2(-c0 / 360 * (1 - 3 * c1 / 56), -1.0 / 6 * (1 - c1 / 10 * (1.0 - c1 / 28)));
        b12 = make_double2(0.5 * (1.0 / 12 * (1 - 2 * c1 / 30 * (1 - 3 * c1 / 112))), 0.5 * (-c0 / 1260 * (1 - c1 / 24)));
        b20 = make_double2(-c0 / 360, -1.0 / 6 * (1 - c1 / 20 * (1 - c1 / 42)));
        b21 = make_double2(1.0 / 24 * (1 - c1 / 15 * (1 - 3 * c1 / 112)), -c0 / 2520);
        b22 = make_double2(0.5 * c0 / 10080, 0.5 * (1.0 / 60 * (1 - c1 / 21 * (1 - c1 / 48))));
    } else {
        int sign = 1;
        if (c0 < 0) {
            sign = -1;
            c0 = -c0;
        }
        const double eps = (c0max - c0) / c0max;
        double theta;
        if (eps < 0) theta = 0.0;
        else if (eps < 1e-3)
            theta = sqrt(2 * eps) *
                    (1 + (1.0 / 12 + (3.0 / 160 + (5.0 / 896 + (35.0 / 18432 + 63.0 / 90112 * eps) * eps) * eps) * eps) * eps);
        else theta = acos(c0 / c0max);
        const double u = sqrt(c1 / 3) * cos(theta / 3);
        const double w = sqrt(c1) * sin(theta / 3);
        const double u2 = u * u, w2 = w * w, u2mw2 = u2 - w2, w2p3u2 = w2 + 3 * u2, w2m3u2 = w2 - 3 * u2;
        const double cu = cos(u), c2u = cos(2 * u), su = sin(u), s2u = sin(2 * u), cw = cos(w);
        double xi0w, xi1w;
        if (fabs(w) < 0.05) {
            const double t0 = w * w, t1 = 1 - t0 / 42, t2 = 1.0 - t0 / 20 * t1;
            xi0w = 1 - t0 / 6 * t2;
            xi1w = -(1 - w2 * (1 - w2 * (1 - w2 / 54) / 28) / 10) / 3;
        } else {
            xi0w = sin(w) / w;
            xi1w = cw / w2 - sin(w) / (w2 * w);
        }
        const double denom = 1 / (9 * u * u - w * w);
        f0 = make_double2((u2mw2 * c2u + cu * 8 * u2 * cw + 2 * su * u * w2p3u2 * xi0w) * denom,
                          (u2mw2 * s2u - su * 8 * u2 * cw + cu * 2 * u * w2p3u2 * xi0w) * denom);
        f1 = make_double2((2 * u * c2u - cu * 2 * u * cw - su * w2m3u2 * xi0w) * denom,
                          (2 * u * s2u + su * 2 * u * cw - cu * w2m3u2 * xi0w) * denom);
        f2 = make_double2((c2u - cu * cw - 3 * su * u * xi0w) * denom, (s2u + su * cw - cu * 3 * u * xi0w) * denom);
        const double2 r01 = make_double2(
            2 * c2u * u + s2u * (-2 * u2 + 2 * w2) + 2 * cu * u * (8 * cw + 3 * u2 * xi0w + w2 * xi0w) +
                su * (-8 * cw * u2 + 18 * u2 * xi0w + 2 * w2 * xi0w),
            -8 * cw * (2 * su * u + cu * u2) + 2 * (s2u * u + c2u * u2 - c2u * w2) +
                2 * (9 * cu * u2 - 3 * su * u * u2 + cu * w2 - su * u * w2) * xi0w);
        const double2 r11 = make_double2(2 * c2u - 4 * s2u * u + su * (2 * cw * u + 6 * u * xi0w) +
                                             cu * (-2 * cw + 3 * u2 * xi0w - w2 * xi0w),
                                         2 * s2u + 4 * c2u * u + 2 * cw * (su + cu * u) + (6 * cu * u - 3 * su * u2 + su * w2) * xi0w);
        const double2 r21 = make_double2(-2 * s2u + cw * su - 3 * (su + cu * u) * xi0w,
                                         2 * c2u + cu * cw + (-3 * cu + 3 * su * u) * xi0w);
        const double2 r02 = make_double2(-2 * c2u + 2 * cw * su * u + 2 * su * u * xi0w - 8 * cu * u2 * xi0w +
                                             6 * su * u * u2 * xi1w,
                                         2 * (-s2u + 4 * su * u2 * xi0w + cu * u * (cw + xi0w + 3 * u2 * xi1w)));
        const double2 r12 = make_double2(2 * cu * u * xi0w + su * (-cw - xi0w + 3 * u2 * xi1w),
                                         -2 * su * u * xi0w - cu * (cw + xi0w - 3 * u2 * xi1w));
        const double2 r22 = make_double2(cu * xi0w - 3 * su * u * xi1w, -(su * xi0w) - 3 * cu * u * xi1w);
        const double h = 0.5 * denom * denom;
        const double a1 = 2.0 * u, a2 = 3.0 * u * u - w * w, a3 = -2.0 * (15.0 * u * u + w * w);
        auto B1 = [&](double2 r1, double2 r2, double2 f) {
            return make_double2(h * (a1 * r1.x + a2 * r2.x + a3 * f.x), h * (a1 * r1.y + a2 * r2.y + a3 * f.y));
        };
        auto B2 = [&](double2 r1, double2 r2, double2 f) {
            return make_double2(h * (r1.x - 3.0 * u * r2.x - 24.0 * u * f.x), h * (r1.y - 3.0 * u * r2.y - 24.0 * u * f.y));
        };
        b10 = B1(r01, r02, f0);
        b11 = B1(r11, r12, f1);
        b12 = B1(r21, r22, f2);
        b20 = B2(r01, r02, f0);
        b21 = B2(r11, r12, f1);
        b22 = B2(r21, r22, f2);
        if (sign == -1) {
            b10 = cconj(b10);
            b11 = make_double2(-b11.x, b11.y);
            b12 = cconj(b12);
            b20 = make_double2(-b20.x, b20.y);
            b21 = cconj(b21);
            b22 = make_double2(-b22.x, b22.y);
            f0 = cconj(f0);
            f1 = make_double2(-f1.x, f1.y);
            f2 = cconj(f2);
        }
    }
    const double2 z = make_double2(0.0, 0.0), one = make_double2(1.0, 0.0);
    Su3 Bm, Q, Q2, T, USP, QU, UQ;
    itamat_poly_no3rdrow(z, one, z, q, Q);
    itamat_poly_row2(z, one, z, q, Q);
    itamat_poly_no3rdrow(z, z, one, q, Q2);
    itamat_poly_row2(z, z, one, q, Q2);
    // tr_k = Tr(Sigma' B_k U)
    double2 tr1 = z, tr2 = z;
    itamat_poly_no3rdrow(b10, b11, b12, q, Bm);
    itamat_poly_row2(b10, b11, b12, q, Bm);
    su3_mul(Bm, U, T);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) tr1 = cfma(SP.c[a][b], T.c[b][a], tr1);
    itamat_poly_no3rdrow(b20, b21, b22, q, Bm);
    itamat_poly_row2(b20, b21, b22, q, Bm);
    su3_mul(Bm, U, T);
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) tr2 = cfma(SP.c[a][b], T.c[b][a], tr2);
    su3_mul(U, SP, USP);
    su3_mul(Q, USP, QU);
    su3_mul(USP, Q, UQ);
    Su3 G;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            double2 s = cmul(tr1, Q.c[a][b]);
            s = cfma(tr2, Q2.c[a][b], s);
            s = cfma(f1, USP.c[a][b], s);
            s = cfma(f2, cadd(QU.c[a][b], UQ.c[a][b]), s);
            G.c[a][b] = s;
        }
    const double third = 0.33333333333333333333333;
    L.rc00 = (2 * G.c[0][0].x - G.c[1][1].x - G.c[2][2].x) * third;
    L.rc11 = (2 * G.c[1][1].x - G.c[0][0].x - G.c[2][2].x) * third;
    L.c01 = cscale(cadd(G.c[0][1], cconj(G.c[1][0])), 0.5);
    L.c02 = cscale(cadd(G.c[0][2], cconj(G.c[2][0])), 0.5);
    L.c12 = cscale(cadd(G.c[1][2], cconj(G.c[2][1])), 0.5);
}

// staples -> QA -> Lambda for every link (compute_sigma_from_sigma_prime ... first half)
__global__ __launch_bounds__(kBlock) void k_qa_lambda(GaugePtrs G, const double2* __restrict__ SP, size_t su3_stride,
                                                     void* __restrict__ qa, void* __restrict__ lam, LinkGeom g,
                                                     double rho, double c0, int qa_given) {
    int dir, idxh;
    if (!link_of_tile(g.sizeh, dir, idxh)) return;
    Su3 U, S;
    load_link(G.at(dir), g.sizeh, idxh, U);
    Tamat q;
    if (qa_given) {
        ta_load(qa, dir, g.sizeh, idxh, q);
    } else {
        int d[4];
        decode(idxh, dir & 1, g, d);
        Su3 C, X;
        staple_sum(G, g, d, idxh, dir >> 1, dir & 1, c0, C);
        su3_mul(U, C, X);
        ta_part(X, rho / c0, q);
        ta_store(qa, dir, g.sizeh, idxh, q);
    }
    load_gl3((const double2*)((const char*)SP + dir * su3_stride), g.sizeh, idxh, S);
    Thmat L;
    lambda_local(q, S, U, L);
    th_store(lam, dir, g.sizeh, idxh, L);
}

// Sigma = Sigma' exp(iQ) + i rho (staple terms with Lambda) (compute_sigma, stouting.c:1175-1302)
__global__ __launch_bounds__(kBlock) void k_sigma(GaugePtrs G, const void* __restrict__ lam,
                                                 const void* __restrict__ qa, double2* __restrict__ S,
                                                 size_t su3_stride, LinkGeom g, double rho) {
    int dir, idxh;
    if (!link_of_tile(g.sizeh, dir, idxh)) return;
    const int mu = dir >> 1, parity = dir & 1, sh = g.sizeh;
    int d[4];
    decode(idxh, parity, g, d);
    Tamat q;
    ta_load(qa, dir, sh, idxh, q);
    Su3 E, SPm, res;
    ch_exp_minus_qa(q, E);
    rebuild3row(E);
    double2* Sd = (double2*)((char*)S + dir * su3_stride);
    load_gl3(Sd, sh, idxh, SPm);
    su3_mul(SPm, E, res);
    Thmat lt;
    Su3 LD;
    th_load(lam, dir, sh, idxh, lt);
    th_full(lt, LD);
    const double2 irho = make_double2(0.0, rho), mirho = make_double2(0.0, -rho);
    const int i_pmu = shifted(d, mu, 1, 0, 0, g);
    for (int it = 0; it < 3; it++) {
        const int nu = c_perp[mu][it];
        const int i_pnu = shifted(d, nu, 1, 0, 0, g);
        Su3 A, B, C, AB, ABC, X, L1, L2;
        // right part: A = U_nu(x+mu), B = U_mu(x+nu)^+, C = U_nu(x)^+
        load_link(G.at(2 * nu + !parity), sh, i_pmu, A);
        load_link(G.at(2 * mu + !parity), sh, i_pnu, B);
        load_link(G.at(2 * nu + parity), sh, idxh, C);
        su3_mul_dag(A, B, AB);
        su3_mul_dag(AB, C, ABC);
        // i rho ABC (L_mu(x) - L_nu(x))
        th_load(lam, 2 * nu + parity, sh, idxh, lt);
        th_full(lt, L1);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) L1.c[a][b] = csub(LD.c[a][b], L1.c[a][b]);
        su3_mul(ABC, L1, X);
        su3_axpy(irho, X, res);
        // i rho L_nu(x+mu) ABC
        th_load(lam, 2 * nu + !parity, sh, i_pmu, lt);
        th_full(lt, L2);
        su3_mul(L2, ABC, X);
        su3_axpy(irho, X, res);
        // -i rho AB L_mu(x+nu) C
        th_load(lam, 2 * mu + !parity, sh, i_pnu, lt);
        th_full(lt, L2);
        su3_mul(AB, L2, X);
        su3_mul_dag(X, C, X);
        su3_axpy(mirho, X, res);
        // left part: A = U_nu(x+mu-nu)^+, B = U_mu(x-nu)^+, C = U_nu(x-nu)
        const int i_mnu = shifted(d, nu, -1, 0, 0, g);
        const int i_pmu_mnu = shifted(d, mu, 1, nu, -1, g);
        load_link(G.at(2 * nu + parity), sh, i_pmu_mnu, A);
        load_link(G.at(2 * mu + !parity), sh, i_mnu, B);
        load_link(G.at(2 * nu + !parity), sh, i_mnu, C);
        su3_mul(B, A, AB);  // (A^+ B^+) = (B A)^+
        // i rho A^+ B^+ (L_nu(x-nu) - L_mu(x-nu)) C
        th_load(lam, 2 * nu + !parity, sh, i_mnu, lt);
        th_full(lt, L1);
        th_load(lam, 2 * mu + !parity, sh, i_mnu, lt);
        th_full(lt, L2);
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) L1.c[a][b] = csub(L1.c[a][b], L2.c[a][b]);
        su3_dag_mul(AB, L1, X);   // (BA)^+ (G - E)
        su3_mul(X, C, X);
        su3_axpy(irho, X, res);
        // i rho A^+ B^+ C L_mu(x)
        su3_dag_mul(AB, C, X);
        su3_mul(X, LD, X);
        su3_axpy(irho, X, res);
        // -i rho A^+ L_nu(x+mu-nu) B^+ C
        th_load(lam, 2 * nu + parity, sh, i_pmu_mnu, lt);
        th_full(lt, L2);
        su3_dag_mul(A, L2, X);
        su3_dag_mul(B, C, L1);
        su3_mul(X, L1, X);
        su3_axpy(mirho, X, res);
    }
    store_rows(Sd, sh, idxh, res, 3);
}

// unitarize_conf (su3_utilities.c:17-40, loc_unitarize_conf su3_utilities.h:68-113)
__global__ __launch_bounds__(kBlock) void k_unitarize(double2* __restrict__ u, size_t su3_stride, int sizeh) {
    const int gid = blockIdx.x * kBlock + (int)threadIdx.x;
    if (gid >= 8 * sizeh) return;
    int dir, i;
    link_of(gid, sizeh, dir, i);
    double2* U = (double2*)((char*)u + dir * su3_stride);
    double2 a0 = U[i], a1 = U[sizeh + i], a2 = U[2 * sizeh + i];
    double2 b0 = U[3 * sizeh + i], b1 = U[4 * sizeh + i], b2 = U[5 * sizeh + i];
    double n = 1.0 / sqrt(cnorm2(a0) + cnorm2(a1) + cnorm2(a2));
    a0 = cscale(a0, n);
    a1 = cscale(a1, n);
    a2 = cscale(a2, n);
    double2 sp = cmulc(a0, b0);
    sp = cfmac(a1, b1, sp);
    sp = cfmac(a2, b2, sp);
    b0 = csub(b0, cmul(sp, a0));
    b1 = csub(b1, cmul(sp, a1));
    b2 = csub(b2, cmul(sp, a2));
    n = 1.0 / sqrt(cnorm2(b0) + cnorm2(b1) + cnorm2(b2));
    U[i] = a0;
    U[sizeh + i] = a1;
    U[2 * sizeh + i] = a2;
    U[3 * sizeh + i] = cscale(b0, n);
    U[4 * sizeh + i] = cscale(b1, n);
    U[5 * sizeh + i] = cscale(b2, n);
}

// ---------------------------------------------------------------- launchers
static LinkGeom link_geom() {
    const Lattice& L = ctx().lat;
    LinkGeom g;
    for (int i = 0; i < 4; i++) g.nd[i] = L.nd[i];
    g.sizeh = (int)L.sizeh;
    return g;
}
static GaugePtrs gauge_ptrs(const void* u) {
    GaugePtrs G;
    const size_t st = ctx().lat.su3_bytes();
    G.base = (const char*)u;
    G.stride = st;
    return G;
}
static int nblk8() { return blocks_for(8 * ctx().lat.sizeh); }
static int nblk_tiles() { return 8 * blocks_for(ctx().lat.sizeh); }  // link_of_tile kernels

bool launch_stout(const void* u, void* uout, double rho, double c0, void* stap, void* qa, void* expo) {
    hipLaunchKernelGGL(k_stout, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(u), (double2*)uout,
                       ctx().lat.su3_bytes(), link_geom(), rho, c0, (double2*)stap, qa, (double2*)expo);
    return hip_ok(hipGetLastError(), "k_stout");
}

bool launch_sigma_chain(void* sigma, void* lam, void* qa, const void* u, double rho, double c0) {
    const size_t st = ctx().lat.su3_bytes();
    hipLaunchKernelGGL(k_qa_lambda, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(u),
                       (const double2*)sigma, st, qa, lam, link_geom(), rho, c0, 0);
    if (!hip_ok(hipGetLastError(), "k_qa_lambda")) return false;
    hipLaunchKernelGGL(k_sigma, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(u), (const void*)lam,
                       (const void*)qa, (double2*)sigma, st, link_geom(), rho);
    return hip_ok(hipGetLastError(), "k_sigma");
}

// tile extent of the plane kernel along a direction of extent n (0: none fits)
static int gf_tile(int n) { return n % 8 == 0 ? 8 : (n % 4 == 0 ? 4 : 0); }
static int g_gf_planes = 1;  // the plane kernels where the extents allow (0: the per-link kernel)

template <int TM, int TN>
static void launch_gf_plane(const void* u, void* stap, void* ipdot, void* mom, int mu, int nu, double fact,
                            int flags) {
    const Lattice& L = ctx().lat;
    int oth = 1;
    for (int k = 0; k < 4; k++)
        if (k != mu && k != nu) oth *= L.nd[k];
    const int nb = (L.nd[mu] / TM) * (L.nd[nu] / TN) * oth;
    hipLaunchKernelGGL((k_gf_plane<TM, TN>), dim3(nb), dim3(4 * TM * TN), 0, ctx().stream, gauge_ptrs(u),
                       (double2*)stap, L.su3_bytes(), ipdot, mom, link_geom(), mu, nu, L.par_temp, g_tlsm ? 1 : 0,
                       g_c_zero, fact, flags);
}

static bool launch_gauge_force_planes(const void* u, void* stap, void* ipdot, void* mom, double fact) {
    const Lattice& L = ctx().lat;
    static const int planes[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    int first[4] = {-1, -1, -1, -1}, last[4] = {-1, -1, -1, -1};
    for (int p = 0; p < 6; p++)
        for (int e = 0; e < 2; e++) {
            const int d = planes[p][e];
            if (first[d] < 0) first[d] = p;
            last[d] = p;
        }
    for (int p = 0; p < 6; p++) {
        const int mu = planes[p][0], nu = planes[p][1];
        const int flags = (first[mu] == p ? 1 : 0) | (last[mu] == p ? 2 : 0) | (first[nu] == p ? 4 : 0) |
                          (last[nu] == p ? 8 : 0);
        const int tm = gf_tile(L.nd[mu]), tn = gf_tile(L.nd[nu]);
        if (tm == 8 && tn == 8) launch_gf_plane<8, 8>(u, stap, ipdot, mom, mu, nu, fact, flags);
        else if (tm == 8) launch_gf_plane<8, 4>(u, stap, ipdot, mom, mu, nu, fact, flags);
        else if (tn == 8) launch_gf_plane<4, 8>(u, stap, ipdot, mom, mu, nu, fact, flags);
        else launch_gf_plane<4, 4>(u, stap, ipdot, mom, mu, nu, fact, flags);
    }
    return hip_ok(hipGetLastError(), "k_gf_plane");
}

bool launch_gauge_force(const void* u, void* stap, void* ipdot, void* mom, double fact) {
    const Lattice& L = ctx().lat;
    // lane offsets within a link array are 32-bit (up to 6 planes of 16 sizeh bytes)
    if ((size_t)96 * L.sizeh >= ((size_t)1 << 32)) return fail("gauge force: lattice too large (sizeh %ld)", (long)L.sizeh);
    // the plane kernels when every extent is a multiple of 4
    if (g_gf_planes && gf_tile(L.nd[0]) && gf_tile(L.nd[1]) && gf_tile(L.nd[2]) && gf_tile(L.nd[3]))
        return launch_gauge_force_planes(u, stap, ipdot, mom, fact);
    // whole blocks per link array (the kernel takes the direction from the block)
    const int nb = 8 * (int)((L.sizeh + kGfLinks - 1) / kGfLinks);
    hipLaunchKernelGGL(k_gauge_force, dim3(nb), dim3(3 * kGfLinks), 0, ctx().stream, gauge_ptrs(u), (double2*)stap,
                       L.su3_bytes(), ipdot, link_geom(), L.par_temp, g_tlsm ? 1 : 0, g_c_zero, mom, fact);
    return hip_ok(hipGetLastError(), "k_gauge_force");
}

// kind 0: plaquettes, 1: rectangles; the sum returned on the host
bool gauge_action_sum(const void* u, int kind, double* out) {
    const Lattice& L = ctx().lat;
    const int nb = blocks_for(2 * L.sizeh);
    SolverCtl c = plain_ctl();
    hipLaunchKernelGGL(k_action_sum, dim3(nb), dim3(kBlock), 0, ctx().stream, gauge_ptrs(u), link_geom(),
                       L.par_temp, kind, c);
    if (!hip_ok(hipGetLastError(), "k_action_sum") || !launch_finalize(kFinStore, c, nb)) return false;
    double v = 0.0;
    if (!hip_ok(hipMemcpyAsync(&v, &ctx().state->sum, sizeof(double), hipMemcpyDeviceToHost, ctx().stream),
                "action readback") ||
        !hip_ok(hipStreamSynchronize(ctx().stream), "action readback"))
        return false;
    *out = v;
    return true;
}

double stout_c_zero() { return g_c_zero; }
bool stout_tlsm() { return g_tlsm; }

}  // namespace ostp

using namespace ostp;

// ------------------------------------------------------------ C entry points
static bool sync(const char* nm) { return hip_ok(hipStreamSynchronize(ctx().stream), nm); }
static double rho_of(int istopo) { return istopo ? gl_topo_rho : gl_stout_rho; }

extern "C" {

int ostp_stout_steps(void) { return g_stout_steps; }

void ostp_set_stout(double rho, int steps, double topo_rho, int topo_steps, int topo_action, int tlsm) {
    gl_stout_rho = rho;
    gl_topo_rho = topo_rho;
    g_stout_steps = steps;
    g_topo_steps = topo_steps;
    g_topo_action = topo_action;
    g_c_zero = tlsm ? 5.0 * 0.33333333333333333333333 : 1.0;
    g_tlsm = tlsm != 0;
}

void stout_isotropic(const su3_soa* u, su3_soa* uprime, su3_soa* local_staples, su3_soa* auxiliary,
                     tamat_soa* tipdot, const int istopo) {
    const char* nm = "stout_isotropic";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    const void* du = resolve(u, gb, nm);
    void* dout = resolve(uprime, gb, nm);
    void* dst = local_staples ? resolve(local_staples, gb, nm) : nullptr;
    void* dex = auxiliary ? resolve(auxiliary, gb, nm) : nullptr;
    void* dqa = tipdot ? resolve(tipdot, tb, nm) : nullptr;
    if (!du || !dout || (local_staples && !dst) || (auxiliary && !dex) || (tipdot && !dqa)) return;
    if (launch_stout(du, dout, rho_of(istopo), g_c_zero, dst, dqa, dex)) sync(nm);
}

void stout_wrapper(const su3_soa* tconf_acc, su3_soa* tstout_conf_acc_arr, const int istopo) {
    const char* nm = "stout_wrapper";
    if (!require_init(nm)) return;
    const int steps = (istopo & g_topo_action) ? g_topo_steps : g_stout_steps;
    if (steps <= 0) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes();
    const void* du = resolve(tconf_acc, gb, nm);
    char* arr = (char*)resolve(tstout_conf_acc_arr, steps * gb, nm);
    if (!du || !arr) return;
    const double rho = rho_of(istopo);
    if (!launch_stout(du, arr, rho, g_c_zero, nullptr, nullptr, nullptr)) return;
    for (int level = 1; level < steps; level++)
        if (!launch_stout(arr + (level - 1) * gb, arr + level * gb, rho, g_c_zero, nullptr, nullptr, nullptr)) return;
    sync(nm);
}

void exp_minus_QA_times_conf(const su3_soa* tu, const tamat_soa* QA, su3_soa* tu_out, su3_soa* exp_aux) {
    const char* nm = "exp_minus_QA_times_conf";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    const void* du = resolve(tu, gb, nm);
    const void* dq = resolve(QA, tb, nm);
    void* dout = resolve(tu_out, gb, nm);
    void* dex = exp_aux ? resolve(exp_aux, gb, nm) : nullptr;
    if (!du || !dq || !dout || (exp_aux && !dex)) return;
    hipLaunchKernelGGL(k_exp_times_conf, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(du), dq,
                       (double2*)dout, (double2*)dex, L.su3_bytes(), link_geom());
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

void calc_loc_staples_nnptrick_all_onlyferms(const su3_soa* u, su3_soa* loc_stap) {
    const char* nm = "calc_loc_staples_nnptrick_all_onlyferms";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    void* ds = resolve(loc_stap, 8 * L.su3_bytes(), nm);
    if (!du || !ds) return;
    hipLaunchKernelGGL(k_staples, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(du), (double2*)ds,
                       L.su3_bytes(), link_geom(), g_c_zero);
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

void RHO_times_conf_times_staples_ta_part(const su3_soa* u, const su3_soa* loc_stap, tamat_soa* tipdot, int istopo) {
    const char* nm = "RHO_times_conf_times_staples_ta_part";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* du = resolve(u, 8 * L.su3_bytes(), nm);
    const void* ds = resolve(loc_stap, 8 * L.su3_bytes(), nm);
    void* dq = resolve(tipdot, (size_t)8 * 64 * L.sizeh, nm);
    if (!du || !ds || !dq) return;
    hipLaunchKernelGGL(k_rho_ta, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(du), (const double2*)ds,
                       L.su3_bytes(), dq, link_geom(), rho_of(istopo), g_c_zero);
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

void compute_lambda(thmat_soa* Lm, const su3_soa* SP, const su3_soa* U, const tamat_soa* QA, su3_soa* TMP) {
    const char* nm = "compute_lambda";
    (void)TMP;
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    void* dl = resolve(Lm, tb, nm);
    const void* ds = resolve(SP, gb, nm);
    const void* du = resolve(U, gb, nm);
    const void* dq = resolve(QA, tb, nm);
    if (!dl || !ds || !du || !dq) return;
    hipLaunchKernelGGL(k_qa_lambda, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(du), (const double2*)ds,
                       L.su3_bytes(), const_cast<void*>(dq), dl, link_geom(), 0.0, g_c_zero, 1);
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

void compute_sigma(const thmat_soa* Lm, const su3_soa* U, su3_soa* S, const tamat_soa* QA, su3_soa* TMP,
                   const int istopo) {
    const char* nm = "compute_sigma";
    (void)TMP;
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    const void* dl = resolve(Lm, tb, nm);
    const void* du = resolve(U, gb, nm);
    void* ds = resolve(S, gb, nm);
    const void* dq = resolve(QA, tb, nm);
    if (!dl || !ds || !du || !dq) return;
    hipLaunchKernelGGL(k_sigma, dim3(nblk_tiles()), dim3(kBlock), 0, ctx().stream, gauge_ptrs(du), dl, dq, (double2*)ds,
                       L.su3_bytes(), link_geom(), rho_of(istopo));
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

void compute_sigma_from_sigma_prime_backinto_sigma_prime(su3_soa* Sigma, thmat_soa* Lambda, tamat_soa* QA,
                                                          const su3_soa* U, su3_soa* TMP, const int istopo) {
    const char* nm = "compute_sigma_from_sigma_prime_backinto_sigma_prime";
    (void)TMP;
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    void* ds = resolve(Sigma, gb, nm);
    void* dl = resolve(Lambda, tb, nm);
    void* dq = resolve(QA, tb, nm);
    const void* du = resolve(U, gb, nm);
    if (!ds || !dl || !dq || !du) return;
    if (launch_sigma_chain(ds, dl, dq, du, rho_of(istopo), g_c_zero)) sync(nm);
}

void unitarize_conf(su3_soa* u) {
    const char* nm = "unitarize_conf";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    void* du = resolve(u, 8 * L.su3_bytes(), nm);
    if (!du) return;
    hipLaunchKernelGGL(k_unitarize, dim3(nblk8()), dim3(kBlock), 0, ctx().stream, (double2*)du, L.su3_bytes(),
                       (int)L.sizeh);
    if (hip_ok(hipGetLastError(), nm)) sync(nm);
}

// ipdot_gauge.c:361-372 (GAUGE_ACTION from ostp_set_stout's tlsm flag)
void calc_ipdot_gauge_soloopenacc(const su3_soa* tconf_acc, su3_soa* local_staples, tamat_soa* tipdot) {
    const char* nm = "calc_ipdot_gauge_soloopenacc";
    if (!require_init(nm)) return;
    const Lattice& L = ctx().lat;
    const void* du = resolve(tconf_acc, 8 * L.su3_bytes(), nm);
    void* ds = resolve(local_staples, 8 * L.su3_bytes(), nm);
    void* dt = resolve(tipdot, 8 * (size_t)64 * L.sizeh, nm);
    if (!du || !ds || !dt) return;
    if (launch_gauge_force(du, ds, dt, nullptr, 0.0)) sync(nm);
}

// su3_measurements.c:191-211: sum over sites and the six planes of Re tr of
// the plaquettes (times the defect coefficients with PAR_TEMP).  The
// reference's scratch arguments are not written.
double calc_plaquette_soloopenacc(const su3_soa* tconf_acc, su3_soa* local_plaqs, dcomplex_soa* tr_local_plaqs) {
    (void)local_plaqs;
    (void)tr_local_plaqs;
    const char* nm = "calc_plaquette_soloopenacc";
    if (!require_init(nm)) return NAN;
    const void* du = resolve(tconf_acc, 8 * ctx().lat.su3_bytes(), nm);
    double v = NAN;
    if (du && !gauge_action_sum(du, 0, &v)) return NAN;
    return v;
}

// rectangles.c:1890-1909: the same for the 2x1 and 1x2 rectangles
double calc_rettangolo_soloopenacc(const su3_soa* tconf_acc, su3_soa* local_plaqs, dcomplex_soa* tr_local_plaqs) {
    (void)local_plaqs;
    (void)tr_local_plaqs;
    const char* nm = "calc_rettangolo_soloopenacc";
    if (!require_init(nm)) return NAN;
    const void* du = resolve(tconf_acc, 8 * ctx().lat.su3_bytes(), nm);
    double v = NAN;
    if (du && !gauge_action_sum(du, 1, &v)) return NAN;
    return v;
}

// device time per launch of the gauge-force kernel (HIP events on the
// library stream), with the fused momentum step when mom is not null
int ostp_set_gauge_force_kernel(int planes) {
    const int prev = g_gf_planes;
    g_gf_planes = planes ? 1 : 0;
    return prev;
}

double ostp_time_gauge_force(const su3_soa* u, su3_soa* stap, tamat_soa* ipdot, thmat_soa* mom, double fact,
                             int reps) {
    const char* nm = "ostp_time_gauge_force";
    if (!require_init(nm)) return -1.0;
    const Lattice& L = ctx().lat;
    const size_t gb = 8 * L.su3_bytes(), tb = (size_t)8 * 64 * L.sizeh;
    const void* du = resolve(u, gb, nm);
    void* ds = stap ? resolve(stap, gb, nm) : nullptr;
    void* di = resolve(ipdot, tb, nm);
    void* dm = mom ? resolve(mom, tb, nm) : nullptr;
    if (!du || (stap && !ds) || !di || (mom && !dm) || reps <= 0) return -1.0;
    hipEvent_t e0, e1;
    if (!hip_ok(hipEventCreate(&e0), "event") || !hip_ok(hipEventCreate(&e1), "event")) return -1.0;
    (void)hipEventRecord(e0, ctx().stream);
    for (int r = 0; r < reps; r++)
        if (!launch_gauge_force(du, ds, di, dm, fact)) return -1.0;
    (void)hipEventRecord(e1, ctx().stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 1000.0 * (double)ms / reps;
}

}  // extern "C"
