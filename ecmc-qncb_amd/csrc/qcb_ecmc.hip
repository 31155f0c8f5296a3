// qcb_ecmc.hip -- event-chain Monte Carlo on the GPU (reference
// src/ecmc/ecmc_mpi_cb.cpp:12-637, src/ecmc/ecmc_mpi_cb.h:56-94).
//
// The reference runs one event chain per MPI rank on its L_core^4 block, with
// the block's "c = 0" faces frozen (src/geometry/GeometryCB.cpp:33-56) and
// random shifts between samples.  Here the whole lattice sits on one device
// and is cut into the same blocks: a link (x, mu) is frozen when
// x_d mod b_d == 0 for some d != mu, a non-frozen link of a block only reads
// links of its own block or frozen ones, so every block's chain runs at once,
// one chain per team of 16 lanes.  Chains keep their LocalChainState between samples
// (sample_persistant) in device memory.  Randomness is Philox4x32-10 on
// per-chain counters (scheme in include/qcb.h), so each chain's trajectory
// depends only on (seed, chain index, its own history).  FP64 links.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/qcb.h"
#include "qcb_host.hpp"
#include "qcb_layout.hpp"
#include "qcb_rng.hpp"
#include "qcb_su3.hpp"

using namespace qcb;

namespace {

// LocalChainState (ecmc_mpi_cb.h:15-35) plus the chain's Philox position.
struct EcmcChainDev {
    double R[18];
    double th_rs, th_rR, par_rs, par_rR;
    unsigned long long events, lifts, revs, set_counter, set_gen, rng_n;
    uint32_t site;
    int mu, eps, initialized, status, pad;
};

struct EcmcArgs {
    double beta, theta_sample, theta_refresh_site, theta_refresh_R, epsilon_set;
    int poisson;
    int b[4], nb[4], L[4];
    uint32_t key0, key1;
    unsigned long long max_events;
    int n_chains;
    uint32_t bvol;   // sites per block
};

constexpr uint64_t kSetSize = 10001;   // ecmc_set (src/su3/utils.cpp:119-127), 1 + 5000 pairs
// A chain is worked by a team of 16 lanes: lanes 0-5 form and solve the six
// staples of an event in parallel, lanes 0-7 the four plaquettes of a lift
// against R and R';
// the rest of the event is computed identically by every lane of the team.
// A wavefront's teams run one instruction stream, so more chains per
// wavefront share its instructions, while fewer leave less to run one after
// the other where they branch apart and let more wavefronts share the SIMDs:
// four chains per wavefront put the 4096 chains of a 32^4 lattice in 4^4
// blocks on all 1024 SIMDs, one wavefront each (5% more events/s than eight
// chains per wavefront on half of the SIMDs).
constexpr int kTeam = 16;
constexpr int kThreads = 64;   // one wavefront per block: 4 chains

__device__ inline int dsgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

__device__ inline void gco(const EcmcArgs& a, uint32_t s, int c[4]) {
    c[0] = (int)(s % (uint32_t)a.L[0]);
    s /= (uint32_t)a.L[0];
    c[1] = (int)(s % (uint32_t)a.L[1]);
    s /= (uint32_t)a.L[1];
    c[2] = (int)(s % (uint32_t)a.L[2]);
    c[3] = (int)(s / (uint32_t)a.L[2]);
}

__device__ inline uint32_t glx(const EcmcArgs& a, const int c[4]) {
    return (uint32_t)c[0] + (uint32_t)a.L[0] * ((uint32_t)c[1] + (uint32_t)a.L[1] * ((uint32_t)c[2] + (uint32_t)a.L[2] * (uint32_t)c[3]));
}

// Inside the event loop a site is carried as its offsets from the chain's
// block origin.  Every site an event touches lies within [0, b_d] in each
// direction: a non-frozen link (x, mu) has offsets 1..b_d-1 for d != mu and
// 0..b_d-1 for d = mu, and its staples and plaquettes reach x+mu, x+-nu and
// x+mu-nu only.  So a neighbour is one add with no wrap, and offset b_d is the
// next block's face (offset 0 there).
struct Site {
    int o[4];
};

__device__ inline Site gsh(Site x, int d1, int s1, int d2 = 0, int s2 = 0) {
#pragma unroll
    for (int d = 0; d < 4; ++d) x.o[d] += (d == d1 ? s1 : 0) + (d == d2 ? s2 : 0);
    return x;
}

__device__ inline bool same_site(const Site& x, const Site& y) {
    return x.o[0] == y.o[0] && x.o[1] == y.o[1] && x.o[2] == y.o[2] && x.o[3] == y.o[3];
}

// frozen: x_d mod b_d == 0 for some d != mu, i.e. an offset of 0 or b_d
__device__ inline bool frozen_near(const EcmcArgs& a, const Site& x, int mu) {
    bool f = false;
#pragma unroll
    for (int d = 0; d < 4; ++d) f = f || (d != mu && (x.o[d] == 0 || x.o[d] == a.b[d]));
    return f;
}

// lexicographic index of a site given by its offsets from origin org
__device__ inline uint32_t lex_of(const EcmcArgs& a, const int org[4], const Site& x) {
    int c[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        c[d] = org[d] + x.o[d];
        c[d] = c[d] >= a.L[d] ? c[d] - a.L[d] : c[d];
    }
    return glx(a, c);
}

// The chains work on a chain-local copy of the links: block after block
// (block index = chain index), a block's sites in lexicographic order, the
// four links of a site, 18 doubles per link.  A link is 144 contiguous bytes
// and a chain's links stay within its block's b^4 * 576 bytes, so the links
// an event touches share cache lines and pages.
__device__ inline size_t blk_link(const EcmcArgs& a, const int bc[4], const Site& x, int mu) {
    uint32_t blk = 0, loc = 0;
#pragma unroll
    for (int d = 3; d >= 0; --d) {
        const bool nx = x.o[d] == a.b[d];
        int bd = bc[d] + (nx ? 1 : 0);
        bd = bd == a.nb[d] ? 0 : bd;
        blk = blk * (uint32_t)a.nb[d] + (uint32_t)bd;
        loc = loc * (uint32_t)a.b[d] + (uint32_t)(nx ? 0 : x.o[d]);
    }
    return (((size_t)blk * a.bvol + loc) * 4 + (size_t)mu) * 18;
}

__device__ inline M3<double> ldl(const double* __restrict__ E, const EcmcArgs& a, const int bc[4], const Site& x,
                                 int mu) {
    const double2* p = (const double2*)(E + blk_link(a, bc, x, mu));
    M3<double> m;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double2 v = p[k];
        m.v[2 * k] = v.x;
        m.v[2 * k + 1] = v.y;
    }
    return m;
}

__device__ inline void stl(double* __restrict__ E, const EcmcArgs& a, const int bc[4], const Site& x, int mu,
                           const M3<double>& m) {
    double2* p = (double2*)(E + blk_link(a, bc, x, mu));
#pragma unroll
    for (int k = 0; k < 9; ++k) p[k] = make_double2(m.v[2 * k], m.v[2 * k + 1]);
}

// Copy the lattice's links (even/odd AoSoA) into the chain-local copy, or
// back; one thread per link, in the copy's order.
__global__ void k_ecmc_pack(const LatDev<double>* __restrict__ lat, double* __restrict__ E, EcmcArgs a, int back) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint32_t)a.n_chains * a.bvol * 4u) return;
    const LatDev<double>& L = *lat;
    const int mu = (int)(t & 3u);
    uint32_t loc = (t >> 2) % a.bvol, blk = (t >> 2) / a.bvol;
    int c[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        c[d] = (int)(blk % (uint32_t)a.nb[d]) * a.b[d] + (int)(loc % (uint32_t)a.b[d]);
        blk /= (uint32_t)a.nb[d];
        loc /= (uint32_t)a.b[d];
    }
    const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
    double2* u = (double2*)L.U + link_base<double>(L.S, mu * 2 + par, glx(a, c) >> 1);
    double2* e = (double2*)(E + (size_t)t * 18);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if (back) u[64 * k] = e[k];
        else e[k] = u[64 * k];
    }
}

__device__ inline M3<double> dag(const M3<double>& a) {
    M3<double> o;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            o.re(i, j) = a.re(j, i);
            o.im(i, j) = -a.im(j, i);
        }
    return o;
}

// solve_reject_fast (ecmc_mpi_cb.h:56-94)
__device__ double reject_fast(double A, double B, double gamma, int eps) {
    B = (eps == -1) ? -B : B;
    const double R = hypot(A, B);
    const double invR = 1.0 / R;
    const double period = 2.0 * R;
    const double discarded = floor(gamma / period);
    gamma -= discarded * period;
    double phi = atan2(-A, B);
    phi += (phi < 0.0) ? (2.0 * M_PI) : 0.0;
    const double p1 = R - A;
    double alpha;
    if (phi < (M_PI * 0.5) || phi > (M_PI * 1.5))
        alpha = (gamma > p1) ? (gamma - p1) * invR - 1.0 : (gamma + A) * invR;
    else
        alpha = gamma * invR - 1.0;
    alpha = (alpha > 1.0) ? 1.0 : ((alpha < -1.0) ? -1.0 : alpha);
    double theta = phi + asin(alpha);
    theta += (theta < 0.0) ? (2.0 * M_PI) : 0.0;
    theta -= (theta >= 2.0 * M_PI) ? (2.0 * M_PI) : 0.0;
    return theta + 2.0 * M_PI * discarded;
}

// compute_ds (ecmc_mpi_cb.cpp:263-275): Im[(R^dag P R)_00 - (R^dag P R)_11];
// only columns 0 and 1 of P R are formed
__device__ double compute_ds(const M3<double>& P, const M3<double>& R) {
    double res = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
#pragma unroll
        for (int col = 0; col < 2; ++col) {
            double mr = P.re(k, 0) * R.re(0, col), mi = P.re(k, 0) * R.im(0, col);
            mr = fma(-P.im(k, 0), R.im(0, col), mr);
            mi = fma(P.im(k, 0), R.re(0, col), mi);
#pragma unroll
            for (int l = 1; l < 3; ++l) {
                mr = fma(P.re(k, l), R.re(l, col), mr);
                mr = fma(-P.im(k, l), R.im(l, col), mr);
                mi = fma(P.re(k, l), R.im(l, col), mi);
                mi = fma(P.im(k, l), R.re(l, col), mi);
            }
            // Im(conj(R_k,col) (P R)_k,col), added for column 0 and subtracted for column 1
            const double sg = col == 0 ? 1.0 : -1.0;
            res = fma(sg * R.re(k, col), mi, res);
            res = fma(-sg * R.im(k, col), mr, res);
        }
    }
    return res;
}

// update (ecmc_mpi_cb.cpp:357-372): U <- R diag(e^{i xi}, e^{-i xi}, 1) R^dag U,
// given T = R^dag U (already formed for the rejection angles)
__device__ M3<double> ecmc_update(M3<double> T, double theta, int eps, const M3<double>& R) {
    const double xi = eps * theta;
    double s, c;
    sincos(xi, &s, &c);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        double a = T.re(0, j), b = T.im(0, j);
        T.re(0, j) = a * c - b * s;
        T.im(0, j) = a * s + b * c;
        a = T.re(1, j);
        b = T.im(1, j);
        T.re(1, j) = a * c + b * s;
        T.im(1, j) = -a * s + b * c;
    }
    return mul_nn(R, T);
}

__device__ inline U4 ctr_ev(uint32_t chain, unsigned long long n, uint32_t call, const EcmcArgs& a) {
    return philox4x32_10(U4{chain, (uint32_t)n, 0xEC000000u | ((uint32_t)(n >> 32) & 0x00FFFFFFu), call}, a.key0,
                         a.key1);
}

// random_SU3_epsilon (src/su3/utils.cpp:66-103): X01 X02 X12 near the identity.
// X_f is the identity outside rows/columns (i, j), so M <- M X_f changes
// columns i and j only; the sums are those of the full product with its
// exact zero terms left out, in the same order.  a0 = sqrt(1 - eps^2), the
// same for every factor, is formed once per launch.
__device__ M3<double> random_su3_eps(double eps, double a0, const double r[9]) {
    M3<double> M;
    m3_identity(M);
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        const int i = f == 2 ? 1 : 0, j = f == 0 ? 1 : 2;
        const double n = sqrt(r[3 * f] * r[3 * f] + r[3 * f + 1] * r[3 * f + 1] + r[3 * f + 2] * r[3 * f + 2]);
        const double a1 = eps * r[3 * f] / n, a2 = eps * r[3 * f + 1] / n,
                     a3 = eps * r[3 * f + 2] / n;
        // X_ii = a0 + i a3, X_jj = a0 - i a3, X_ij = a2 + i a1, X_ji = -a2 + i a1
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            const double mir = M.re(row, i), mii = M.im(row, i), mjr = M.re(row, j), mji = M.im(row, j);
            double cr = mir * a0, ci = mir * a3, dr = mir * a2, di = mir * a1;
            cr = fma(-mii, a3, cr);
            ci = fma(mii, a0, ci);
            dr = fma(-mii, a1, dr);
            di = fma(mii, a2, di);
            cr = fma(mjr, -a2, cr);
            ci = fma(mjr, a1, ci);
            dr = fma(mjr, a0, dr);
            di = fma(mjr, -a3, di);
            cr = fma(-mji, a1, cr);
            ci = fma(mji, -a2, ci);
            dr = fma(-mji, -a3, dr);
            di = fma(mji, a0, di);
            M.re(row, i) = cr;
            M.im(row, i) = ci;
            M.re(row, j) = dr;
            M.im(row, j) = di;
        }
    }
    return M;
}

// element i of the chain's matrix set of generation g: 0 the identity,
// 2m+1 a near-identity matrix from {chain, g, 0xE5000000 | m, 0..2}, 2m+2 its adjoint
__device__ M3<double> set_element(uint32_t chain, unsigned long long g, unsigned long long i, const EcmcArgs& a,
                                  double eps_a0) {
    M3<double> M;
    if (i == 0) {
        m3_identity(M);
        return M;
    }
    const uint32_t m = (uint32_t)((i - 1) / 2);
    double r[9];
#pragma unroll
    for (uint32_t k = 0; k < 3; ++k) {
        U4 o = philox4x32_10(U4{chain, (uint32_t)g, 0xE5000000u | m, k}, a.key0, a.key1);
        r[3 * k] = u32d(o.x) - 0.5;
        r[3 * k + 1] = u32d(o.y) - 0.5;
        r[3 * k + 2] = u32d(o.z) - 0.5;
    }
    M = random_su3_eps(a.epsilon_set, eps_a0, r);
    return (i & 1) ? M : dag(M);
}

// staple j of link (x, mu) (ecmc_mpi_cb.cpp:12-36): nu = j/2-th direction != mu
__device__ inline int dir_of(int mu, int j) {
    const int k = j >> 1;
    return k + (k >= mu ? 1 : 0);
}

// a or a^dag, chosen per lane without a branch
__device__ inline M3<double> cond_dag(const M3<double>& a, bool d) {
    M3<double> o;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            o.re(i, j) = d ? a.re(j, i) : a.re(i, j);
            o.im(i, j) = d ? -a.im(j, i) : a.im(i, j);
        }
    return o;
}

// Upper staple (j even) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag = [C B A^dag]^dag
// with A = U_nu(x+mu), B = U_mu(x+nu), C = U_nu(x); lower staple (j odd)
// U_nu(x+mu-nu)^dag U_mu(x-nu)^dag U_nu(x-nu) = [C^dag B A]^dag with
// A = U_nu(x+mu-nu), B = U_mu(x-nu), C = U_nu(x-nu).  Both are the same two
// products of per-lane operands, so the six lanes of a team follow one path.
__device__ inline void staple_links(const double* __restrict__ E, const EcmcArgs& a, const int bc[4], const Site& x,
                                    int mu, int j, M3<double>& A, M3<double>& B, M3<double>& C) {
    const int nu = dir_of(mu, j);
    const bool up = (j & 1) == 0;
    const Site xb = gsh(x, nu, up ? 1 : -1);
    const Site xa = gsh(x, mu, 1, nu, up ? 0 : -1);
    const Site xc = up ? x : xb;
    A = ldl(E, a, bc, xa, nu);
    B = ldl(E, a, bc, xb, mu);
    C = ldl(E, a, bc, xc, nu);
}

__device__ inline M3<double> staple_of(bool up, const M3<double>& A, const M3<double>& B, const M3<double>& C) {
    return dag(mul_nn(mul_nn(cond_dag(C, !up), B), cond_dag(A, up)));
}

// the four links of plaquette j, the current link first (GeometryCB.cpp:60-90)
__device__ inline void plaq_links(const Site& x, int mu, int j, Site s[4], int d[4]) {
    const int nu = dir_of(mu, j);
    s[0] = x;
    d[0] = mu;
    if ((j & 1) == 0) {
        s[1] = gsh(x, mu, 1); d[1] = nu;
        s[2] = gsh(x, nu, 1); d[2] = mu;
        s[3] = x;                d[3] = nu;
    } else {
        s[1] = gsh(x, mu, 1, nu, -1); d[1] = nu;
        s[2] = gsh(x, nu, -1);        d[2] = mu;
        s[3] = s[2];                     d[3] = nu;
    }
}

// The plaquettes P_i of lift_improved_fast (ecmc_mpi_cb.cpp:296-318) are the
// cyclic readings of one loop M = P_0 or of its inverse, each starting at
// link i: P_i = V_i^dag M^(s_i) V_i with V_i unitary and M^(-) = M^dag.
//   forward  (j even): M = U0 U1 U2^dag U3^dag,  V = 1, U0, U3, 1,
//                      s = +, +, -, -
//   backward (j odd):  M = U0 U1^dag U2^dag U3,  V = 1, U3^dag U2, U3^dag, U3^dag,
//                      s = +, -, -, +
// (e.g. forward P_2 = U3^dag M^dag U3 = U2 U1^dag U0^dag U3).  Since
// (VR)^dag M^dag (VR) is the adjoint of (VR)^dag M (VR), compute_ds(P_i, R) =
// s_i compute_ds(M, V_i R): every lane forms the same M and only its V_i
// differs, so the four lanes follow one path.
__device__ M3<double> plaq_loop(bool fwd, const M3<double>& U0, const M3<double>& U1, const M3<double>& U2,
                                const M3<double>& U3) {
    return mul_nn(mul_nd(mul_nn(U0, cond_dag(U1, !fwd)), U2), cond_dag(U3, fwd));
}

// columns 0 and 1 of V R (all that compute_ds reads of its second argument)
__device__ M3<double> mul_cols01(const M3<double>& V, const M3<double>& R) {
    M3<double> o;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double re = V.re(i, 0) * R.re(0, j), im = V.re(i, 0) * R.im(0, j);
            re = fma(-V.im(i, 0), R.im(0, j), re);
            im = fma(V.im(i, 0), R.re(0, j), im);
#pragma unroll
            for (int k = 1; k < 3; ++k) {
                re = fma(V.re(i, k), R.re(k, j), re);
                re = fma(-V.im(i, k), R.im(k, j), re);
                im = fma(V.re(i, k), R.im(k, j), im);
                im = fma(V.im(i, k), R.re(k, j), im);
            }
            o.re(i, j) = re;
            o.im(i, j) = im;
        }
        o.re(i, 2) = 0.0;
        o.im(i, 2) = 0.0;
    }
    return o;
}

// orders LDS traffic between the lanes of one wavefront (a block is one
// wavefront, so a team's lanes never span two)
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kThreads) void k_ecmc_sample(double* __restrict__ E, EcmcChainDev* __restrict__ chains,
                                                          EcmcArgs a) {
    const int tid = blockIdx.x * kThreads + threadIdx.x;
    const int cid = tid / kTeam, lane = tid % kTeam;
    if (cid >= a.n_chains) return;   // whole teams leave together
    const uint32_t chain = (uint32_t)cid;
    // the event's proposed R' waits in LDS (one copy per team) from the start
    // of the event to the lift, instead of holding 36 VGPRs through the staples
    __shared__ double2 rn_lds[kThreads / kTeam][9];
    double2* const my_rn = rn_lds[threadIdx.x / kTeam];
    EcmcChainDev st = chains[cid];
    int org[4], bc[4];   // block origin and block coordinates (block index = chain index)
    {
        int r = cid;
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            bc[d] = r % a.nb[d];
            org[d] = bc[d] * a.b[d];
            r /= a.nb[d];
        }
    }
    M3<double> R;
#pragma unroll
    for (int k = 0; k < 18; ++k) R.v[k] = st.R[k];
    Site xs;   // the chain's current site
    {
        int c[4];
        gco(a, st.site, c);
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            const int r = c[d] - org[d];
            xs.o[d] = r < 0 ? r + a.L[d] : r;
        }
    }
    auto draw_site = [&](unsigned long long n) {
        const U4 o3 = ctr_ev(chain, n, 3, a), o4 = ctr_ev(chain, n, 4, a);
        const uint32_t w[4] = {o3.x, o3.y, o3.z, o3.w};
#pragma unroll
        for (int d = 0; d < 4; ++d) xs.o[d] = 1 + (int)__umulhi(w[d], (uint32_t)(a.b[d] - 2));
        st.mu = (int)(o4.x >> 30);
        st.eps = 2 * (int)(o4.y >> 31) - 1;
        return o4;
    };
    if (!st.initialized) {   // sample_persistant's first call (ecmc_mpi_cb.cpp:504-518)
        const U4 o4 = draw_site(0);
        const U4 o5 = ctr_ev(chain, 0, 5, a);
        R = haar_su3_philox(a.key0, a.key1, chain, 0u, 0xED000000u);
        st.th_rs = a.poisson ? -a.theta_refresh_site * log(u53(o4.z, o4.w)) : a.theta_refresh_site;
        st.th_rR = a.poisson ? -a.theta_refresh_R * log(u53(o5.x, o5.y)) : a.theta_refresh_R;
        st.par_rs = st.par_rR = 0.0;
        st.set_counter = st.events = st.lifts = 0;
        st.initialized = 1;
    }
    unsigned long long n = st.rng_n + 1;   // theta_sample of this call
    double th_sample;
    {
        const U4 o6 = ctr_ev(chain, n, 6, a);
        th_sample = a.poisson ? -a.theta_sample * log(u53(o6.x, o6.y)) : a.theta_sample;
    }
    unsigned long long gen = st.set_gen + 1;   // the set of this call
    const double beta_red = -(a.beta / 3.0);
    const double eps_a0 = sqrt(1 - a.epsilon_set * a.epsilon_set);   // a0 of every near-identity factor
    double par_sample = 0.0;
    // status is sticky: a chain that stopped early in any launch keeps its
    // non-zero status until qcb_ecmc_get_chains reports it (or set_chains)
    for (unsigned long long guard = 0;; ++guard) {
        if (guard >= a.max_events) {
            st.status = 1;
            break;
        }
        if (st.set_counter % 10000 == 0) {
            gen += 1;
            st.set_counter = 0;
        }
        n += 1;
        // the event's link loads come first, for every lane (lanes past 5
        // repeat staple 5's addresses and discard the result), so that their
        // latency overlaps the random numbers and R' below
        M3<double> SA, SB, SC;
        staple_links(E, a, bc, xs, st.mu, lane < 6 ? lane : 5, SA, SB, SC);
        const M3<double> Ucur = ldl(E, a, bc, xs, st.mu);
        const U4 o0 = ctr_ev(chain, n, 0, a), o1 = ctr_ev(chain, n, 1, a);
        // the lift's proposed R' depends on neither the links nor the chosen
        // plaquette: forming it here lets it overlap the event's link loads
        // (it is used only if the event is a lift, where R is still this R)
        const U4 o2 = ctr_ev(chain, n, 2, a);
        const unsigned long long iset = __umul64hi(((unsigned long long)o2.x << 32) | o2.y, kSetSize);
        {
            const M3<double> Rn = mul_nn(set_element(chain, gen, iset, a, eps_a0), R);
            wave_sync();   // the previous event's reads are done
            if (lane == 0) {
#pragma unroll
                for (int k = 0; k < 9; ++k) my_rn[k] = make_double2(Rn.v[2 * k], Rn.v[2 * k + 1]);
            }
        }
        // compute_reject_angles_fast (ecmc_mpi_cb.cpp:203-250)
        const M3<double> T = mul_dn(R, Ucur);
        // lane j < 6 solves staple j; the team takes the smallest angle, the
        // lowest j on ties (the sequential scan's first minimum)
        double theta_reject = __builtin_inf();
        int jmin = lane;
        if (lane < 6) {
            const int j = lane;
            const M3<double> S = staple_of((j & 1) == 0, SA, SB, SC);
            // (T S R)_00 and (T S R)_11 from rows 0, 1 of T S and columns 0, 1 of R
            double p00r = 0, p00i = 0, p11r = 0, p11i = 0;
#pragma unroll
            for (int col = 0; col < 3; ++col) {
                double m0r = T.re(0, 0) * S.re(0, col), m0i = T.re(0, 0) * S.im(0, col);
                double m1r = T.re(1, 0) * S.re(0, col), m1i = T.re(1, 0) * S.im(0, col);
                m0r = fma(-T.im(0, 0), S.im(0, col), m0r);
                m0i = fma(T.im(0, 0), S.re(0, col), m0i);
                m1r = fma(-T.im(1, 0), S.im(0, col), m1r);
                m1i = fma(T.im(1, 0), S.re(0, col), m1i);
#pragma unroll
                for (int k = 1; k < 3; ++k) {
                    m0r = fma(T.re(0, k), S.re(k, col), m0r);
                    m0r = fma(-T.im(0, k), S.im(k, col), m0r);
                    m0i = fma(T.re(0, k), S.im(k, col), m0i);
                    m0i = fma(T.im(0, k), S.re(k, col), m0i);
                    m1r = fma(T.re(1, k), S.re(k, col), m1r);
                    m1r = fma(-T.im(1, k), S.im(k, col), m1r);
                    m1i = fma(T.re(1, k), S.im(k, col), m1i);
                    m1i = fma(T.im(1, k), S.re(k, col), m1i);
                }
                p00r = fma(m0r, R.re(col, 0), p00r);
                p00r = fma(-m0i, R.im(col, 0), p00r);
                p00i = fma(m0r, R.im(col, 0), p00i);
                p00i = fma(m0i, R.re(col, 0), p00i);
                p11r = fma(m1r, R.re(col, 1), p11r);
                p11r = fma(-m1i, R.im(col, 1), p11r);
                p11i = fma(m1r, R.im(col, 1), p11i);
                p11i = fma(m1i, R.re(col, 1), p11i);
            }
            // gamma_j = -log u_j: lane j takes only its own exponential draw
            const uint32_t wj = j == 0 ? o0.x : j == 1 ? o0.y : j == 2 ? o0.z : j == 3 ? o0.w : j == 4 ? o1.x : o1.y;
            const double gj = -log(u32d(wj));
            theta_reject = reject_fast((p00r + p11r) * beta_red, (p11i - p00i) * beta_red, gj, st.eps);
            // a NaN angle must win the reduction so the team stops
            if (!(theta_reject >= 0.0)) theta_reject = -1.0;
        }
#pragma unroll
        for (int off = kTeam / 2; off > 0; off >>= 1) {
            const double ov = __shfl_xor(theta_reject, off, kTeam);
            const int oj = __shfl_xor(jmin, off, kTeam);
            if (ov < theta_reject || (ov == theta_reject && oj < jmin)) {
                theta_reject = ov;
                jmin = oj;
            }
        }
        if (!(theta_reject >= 0.0)) {   // a non-finite angle would never reach the sample event
            st.status = 2;
            break;
        }
        const double d_sample = th_sample - par_sample;
        const double d_rs = st.th_rs - st.par_rs;
        const double d_rR = st.th_rR - st.par_rR;
        double step = theta_reject;
        step = d_sample < step ? d_sample : step;
        step = d_rs < step ? d_rs : step;
        step = d_rR < step ? d_rR : step;
        // every event kind moves the current link by the step angle first
        // (ecmc_mpi_cb.cpp:571-628); the kinds differ only in their bookkeeping
        const M3<double> U0 = ecmc_update(T, step, st.eps, R);
        stl(E, a, bc, xs, st.mu, U0);
        st.events++;
        if (step == d_sample) {   // SAMPLE: the chain stops here until the next call
            st.par_rs += d_sample;
            st.par_rR += d_sample;
            break;
        } else if (step == d_rs) {   // REFRESH SITE
            par_sample += d_rs;
            st.par_rR += d_rs;
            st.par_rs = 0.0;
            const U4 o4 = draw_site(n);
            if (a.poisson) st.th_rs = -a.theta_refresh_site * log(u53(o4.z, o4.w));
        } else if (step == d_rR) {   // REFRESH R
            par_sample += d_rR;
            st.par_rs += d_rR;
            st.par_rR = 0.0;
            const U4 o5 = ctr_ev(chain, n, 5, a);
            if (a.poisson) st.th_rR = -a.theta_refresh_R * log(u53(o5.x, o5.y));
            R = haar_su3_philox(a.key0, a.key1, chain, (uint32_t)n, 0xED000000u | ((uint32_t)(n >> 32) & 0x00FFFFFFu));
        } else {   // LIFT (lift_improved_fast, ecmc_mpi_cb.cpp:277-354)
            par_sample += theta_reject;
            st.par_rs += theta_reject;
            st.par_rR += theta_reject;
            Site ls[4];
            int ld[4];
            plaq_links(xs, st.mu, jmin, ls, ld);
            const M3<double> U1 = ldl(E, a, bc, ls[1], ld[1]), U2 = ldl(E, a, bc, ls[2], ld[2]),
                             U3 = ldl(E, a, bc, ls[3], ld[3]);
            const bool fwd = (jmin & 1) == 0;
            // R' from LDS where it is used (each read behind its own fence, so
            // that it is not hoisted back to the start of the lift)
            auto load_rn = [&]() {
                wave_sync();
                M3<double> m;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    const double2 v = my_rn[k];
                    m.v[2 * k] = v.x;
                    m.v[2 * k + 1] = v.y;
                }
                return m;
            };
            // lane i < 4 evaluates plaquette i against R, lane 4 + i against R';
            // every lane then reads all four dS and the chosen dS'
            double my_ds = 0.0;
            if (lane < 8) {
                const int i = lane & 3;
                const bool prime = lane >= 4;
                const Site si = i == 0 ? ls[0] : i == 1 ? ls[1] : i == 2 ? ls[2] : ls[3];
                const int di = i == 0 ? ld[0] : i == 1 ? ld[1] : i == 2 ? ld[2] : ld[3];
                const M3<double> M = plaq_loop(fwd, U0, U1, U2, U3);
                const M3<double> U3d = dag(U3), Vb = mul_dn(U3, U2);
                // V_i: 0 the identity, 1 U0, 2 U3, 3 U3^dag, 4 U3^dag U2
                const int vk = fwd ? (i == 1 ? 1 : (i == 2 ? 2 : 0)) : (i == 0 ? 0 : (i == 1 ? 4 : 3));
                const double sg = fwd ? (i < 2 ? 1.0 : -1.0) : ((i == 1 || i == 2) ? -1.0 : 1.0);
                M3<double> V;
#pragma unroll
                for (int k = 0; k < 18; ++k) {
                    const double id = (k == 0 || k == 8 || k == 16) ? 1.0 : 0.0;   // re of the diagonal
                    V.v[k] = vk == 1 ? U0.v[k] : (vk == 2 ? U3.v[k] : (vk == 3 ? U3d.v[k] : (vk == 4 ? Vb.v[k] : id)));
                }
                const M3<double> Rp = load_rn();
                M3<double> Rs;
#pragma unroll
                for (int k = 0; k < 18; ++k) Rs.v[k] = prime ? Rp.v[k] : R.v[k];
                const double d = sg * compute_ds(M, mul_cols01(V, Rs));
                my_ds = (!prime && frozen_near(a, si, di)) ? 0.0 : d;
            }
            const int base = (threadIdx.x / kTeam) * kTeam;
            double absd[4], sum = 0.0;
            int sgn[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double ds = __shfl(my_ds, base + i, 64);
                sgn[i] = dsgn(ds);
                absd[i] = fabs(ds);
                sum += absd[i];
            }
            double prob[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) prob[i] = absd[i] / sum;
            const double u = u32d(o1.z);   // selectVariable (ecmc_mpi_cb.cpp:252-260)
            const int idx = u < prob[0] ? 0 : (u < prob[0] + prob[1] ? 1 : (u < prob[0] + prob[1] + prob[2] ? 2 : 3));
            const double dS = sgn[idx] * absd[idx];
            const double dSn = __shfl(my_ds, base + 4 + idx, 64);
            double neps = -sgn[idx];
            if (fabs(dS) < fabs(dSn) || u32d(o1.w) < fabs(dSn) / fabs(dS)) {
                R = load_rn();
                neps = -dsgn(dSn);
            }
            const Site xn = idx == 0 ? ls[0] : idx == 1 ? ls[1] : idx == 2 ? ls[2] : ls[3];
            st.revs += same_site(xn, xs) ? 1 : 0;
            st.set_counter++;
            st.lifts++;
            xs = xn;
            st.mu = ld[idx];
            st.eps = (int)neps;
        }
    }
    st.site = lex_of(a, org, xs);
    st.rng_n = n;
    st.set_gen = gen;
#pragma unroll
    for (int k = 0; k < 18; ++k) st.R[k] = R.v[k];
    if (lane == 0) chains[cid] = st;
}

}  // namespace

// ------------------------------------------------------------ C-ABI -----

struct qcb_ecmc {
    qcb_ctx* ctx = nullptr;
    qcb_lattice* lat = nullptr;   // counted in lat->n_users, so it cannot be destroyed first
    EcmcArgs args{};
    int n_chains = 0;
    EcmcChainDev* d_chains = nullptr;
    double* d_links = nullptr;   // chain-local copy of the lattice's links (FP64, 72 B x 4 V)
};

extern "C" int qcb_ecmc_create(qcb_lattice* lat, const qcb_ecmc_params* p, qcb_ecmc** out) {
    if (!lat || !p || !out) return QCB_EINVAL;
    *out = nullptr;
    qcb_ctx* ctx = lat->ctx;
    if (lat->prec != QCB_F64) return qcb_set_error(ctx, QCB_EINVAL, "event chains run on FP64 lattices");
    if (!(p->beta > 0) || !(p->theta_sample > 0) || !(p->theta_refresh_site > 0) || !(p->theta_refresh_R > 0) ||
        !(p->epsilon_set >= 0 && p->epsilon_set <= 1))
        return qcb_set_error(ctx, QCB_EINVAL, "ECMC parameters out of range");
    EcmcArgs a{};
    int n = 1;
    for (int d = 0; d < 4; ++d) {
        if (p->block[d] < 4 || lat->L[d] % p->block[d])
            return qcb_set_error(ctx, QCB_EINVAL, "block extent %d must be >= 4 and divide %d", p->block[d], lat->L[d]);
        a.b[d] = p->block[d];
        a.nb[d] = lat->L[d] / p->block[d];
        a.L[d] = lat->L[d];
        n *= a.nb[d];
    }
    a.beta = p->beta;
    a.theta_sample = p->theta_sample;
    a.theta_refresh_site = p->theta_refresh_site;
    a.theta_refresh_R = p->theta_refresh_R;
    a.epsilon_set = p->epsilon_set;
    a.poisson = p->poisson ? 1 : 0;
    a.key0 = (uint32_t)p->seed;
    a.key1 = (uint32_t)(p->seed >> 32);
    a.max_events = p->max_events ? p->max_events : 1000000000ull;
    a.n_chains = n;
    a.bvol = (uint32_t)(a.b[0] * a.b[1] * a.b[2] * a.b[3]);
    qcb_ecmc* e = new qcb_ecmc();
    e->ctx = ctx;
    e->lat = lat;
    e->args = a;
    e->n_chains = n;
    const size_t link_bytes = (size_t)lat->V * 4 * 18 * sizeof(double);
    if (hipMalloc((void**)&e->d_chains, sizeof(EcmcChainDev) * n) != hipSuccess ||
        hipMemset(e->d_chains, 0, sizeof(EcmcChainDev) * n) != hipSuccess ||
        hipMalloc((void**)&e->d_links, link_bytes) != hipSuccess) {
        if (e->d_chains) (void)hipFree(e->d_chains);
        if (e->d_links) (void)hipFree(e->d_links);
        delete e;
        return qcb_set_error(ctx, QCB_ENOMEM, "cannot allocate %d chain states and a %zu-byte link copy", n,
                             link_bytes);
    }
    ctx->n_live++;
    lat->n_users++;
    *out = e;
    return QCB_OK;
}

extern "C" int qcb_ecmc_destroy(qcb_ecmc* e) {
    if (!e) return QCB_EINVAL;
    qcb_ctx* ctx = e->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(e->d_chains);
    (void)hipFree(e->d_links);
    e->lat->n_users--;
    ctx->n_live--;
    delete e;
    return QCB_OK;
}

extern "C" int qcb_ecmc_n_chains(qcb_ecmc* e, int* n) {
    if (!e || !n) return QCB_EINVAL;
    *n = e->n_chains;
    return QCB_OK;
}

extern "C" int qcb_ecmc_sample(qcb_ecmc* e, int n_samples) {
    if (!e || n_samples < 0) return QCB_EINVAL;
    qcb_lattice* lat = e->lat;
    qcb_ctx* ctx = e->ctx;
    if (n_samples == 0) return QCB_OK;
    const int blocks = (e->n_chains * kTeam + kThreads - 1) / kThreads;
    const uint32_t n_links = (uint32_t)lat->V * 4u;
    const dim3 pack_grid((n_links + 255u) / 256u);
    const LatDev<double>* desc = (const LatDev<double>*)lat->d_desc;
    hipLaunchKernelGGL(k_ecmc_pack, pack_grid, dim3(256), 0, ctx->stream, desc, e->d_links, e->args, 0);
    for (int s = 0; s < n_samples; ++s)
        hipLaunchKernelGGL(k_ecmc_sample, dim3(blocks), dim3(kThreads), 0, ctx->stream, e->d_links, e->d_chains,
                           e->args);
    hipLaunchKernelGGL(k_ecmc_pack, pack_grid, dim3(256), 0, ctx->stream, desc, e->d_links, e->args, 1);
    QCB_HIP(ctx, hipGetLastError());
    return QCB_OK;
}

extern "C" int qcb_ecmc_get_chains(qcb_ecmc* e, double* out) {
    if (!e || !out) return QCB_EINVAL;
    qcb_ctx* ctx = e->ctx;
    std::vector<EcmcChainDev> h(e->n_chains);
    QCB_HIP(ctx, hipMemcpyAsync(h.data(), e->d_chains, sizeof(EcmcChainDev) * h.size(), hipMemcpyDeviceToHost,
                                ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int bad = 0;
    for (int c = 0; c < e->n_chains; ++c) {
        const EcmcChainDev& s = h[c];
        double* o = out + (size_t)c * QCB_ECMC_CHAIN_DOUBLES;
        for (int k = 0; k < 18; ++k) o[k] = s.R[k];
        o[18] = s.th_rs; o[19] = s.th_rR; o[20] = s.par_rs; o[21] = s.par_rR;
        o[22] = (double)s.events; o[23] = (double)s.lifts; o[24] = (double)s.revs;
        o[25] = (double)s.set_counter; o[26] = (double)s.set_gen; o[27] = (double)s.rng_n;
        o[28] = (double)s.site; o[29] = s.mu; o[30] = s.eps; o[31] = s.initialized;
        bad += s.status != 0;
    }
    if (bad) {
        // reported once: clear the statuses so the next call sees only new stops
        for (EcmcChainDev& s : h) s.status = 0;
        QCB_HIP(ctx, hipMemcpyAsync(e->d_chains, h.data(), sizeof(EcmcChainDev) * h.size(), hipMemcpyHostToDevice,
                                    ctx->stream));
        QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
        return qcb_set_error(ctx, QCB_EBUDGET,
                             "%d chains stopped before their sample event since the last report (max_events "
                             "reached or a non-finite rejection angle)", bad);
    }
    return QCB_OK;
}

extern "C" int qcb_ecmc_set_chains(qcb_ecmc* e, const double* in) {
    if (!e || !in) return QCB_EINVAL;
    qcb_ctx* ctx = e->ctx;
    const EcmcArgs& a = e->args;
    std::vector<EcmcChainDev> h(e->n_chains);
    for (int c = 0; c < e->n_chains; ++c) {
        const double* v = in + (size_t)c * QCB_ECMC_CHAIN_DOUBLES;
        // counters and the site are whole numbers carried as doubles
        auto whole = [](double x) { return x >= 0 && x < 9007199254740992.0 && x == (double)(uint64_t)x; };
        for (int k = 22; k <= 28; ++k)
            if (!whole(v[k])) return qcb_set_error(ctx, QCB_EINVAL, "chain %d: field %d is not a count", c, k);
        EcmcChainDev& s = h[c];
        std::memset(&s, 0, sizeof s);
        for (int k = 0; k < 18; ++k) s.R[k] = v[k];
        s.th_rs = v[18]; s.th_rR = v[19]; s.par_rs = v[20]; s.par_rR = v[21];
        s.events = (unsigned long long)v[22]; s.lifts = (unsigned long long)v[23]; s.revs = (unsigned long long)v[24];
        s.set_counter = (unsigned long long)v[25]; s.set_gen = (unsigned long long)v[26];
        s.rng_n = (unsigned long long)v[27];
        s.initialized = v[31] != 0.0;
        // the site must lie in the chain's own block (or be unset for a fresh chain)
        const double site = v[28];
        const uint64_t V = (uint64_t)a.L[0] * a.L[1] * a.L[2] * a.L[3];
        if (s.initialized) {
            if (!(site < (double)V) || !(v[29] == 0 || v[29] == 1 || v[29] == 2 || v[29] == 3) ||
                !(v[30] == 1.0 || v[30] == -1.0))
                return qcb_set_error(ctx, QCB_EINVAL, "chain %d: site/mu/epsilon out of range", c);
            s.mu = (int)v[29];
            s.eps = (int)v[30];
            uint32_t x = (uint32_t)site;
            int r = c;
            for (int d = 0; d < 4; ++d) {
                const int bc = r % a.nb[d];
                r /= a.nb[d];
                const int xd = (int)(x % (uint32_t)a.L[d]);
                x /= (uint32_t)a.L[d];
                if (xd / a.b[d] != bc)
                    return qcb_set_error(ctx, QCB_EINVAL, "chain %d: site outside its block", c);
                // the active link must be one the chain may move (not on a frozen face)
                if (d != s.mu && xd % a.b[d] == 0)
                    return qcb_set_error(ctx, QCB_EINVAL, "chain %d: active link is frozen", c);
            }
            s.site = (uint32_t)site;
        }
    }
    QCB_HIP(ctx, hipMemcpyAsync(e->d_chains, h.data(), sizeof(EcmcChainDev) * h.size(), hipMemcpyHostToDevice,
                                ctx->stream));
    QCB_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return QCB_OK;
}
