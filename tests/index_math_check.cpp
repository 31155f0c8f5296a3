// Host build of the device index arithmetic (qcb_layout.hpp / qcb_host.hpp)
// against plain integer arithmetic: closed-form periodic neighbour steps
// (nb_cb), multiply-high division (fdiv) and the tiled work order
// (tiled_coords).  Built and run by tests/test_index_math.py.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdint>
__attribute__((host)) static inline unsigned int __umulhi(unsigned int a, unsigned int b) {
    return (unsigned int)(((uint64_t)a * b) >> 32);
}
__attribute__((host)) static inline unsigned int __umul24(unsigned int a, unsigned int b) {
    return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
}
#undef __device__
#define __device__ __attribute__((host)) __attribute__((device))
#include "qcb_host.hpp"
#include <cstdio>
#include <random>
using namespace qcb;

static uint32_t plain_nb(const LatDev<float>& L, Coords c, int d1, int s1, int d2, int s2) {
    int v[4] = {c.x, c.y, c.z, c.t};
    for (int d = 0; d < 4; ++d) {
        if (d == d1 && s1) v[d] = (v[d] + s1 + L.L[d]) % L.L[d];
        if (d == d2 && s2) v[d] = (v[d] + s2 + L.L[d]) % L.L[d];
    }
    return (uint32_t)(v[0] >> 1) +
           (uint32_t)L.Lxh * ((uint32_t)v[1] + (uint32_t)L.L[1] * ((uint32_t)v[2] + (uint32_t)L.L[2] * (uint32_t)v[3]));
}

static Coords plain_tiled(const LatDev<float>& L, uint32_t q, int p) {
    uint32_t tau = q / L.tvol, r = q - tau * L.tvol;
    uint32_t r1 = r / L.Lxh, r2 = r1 / L.TL[1], r3 = r2 / L.TL[2];
    int xh = r - r1 * L.Lxh, yl = r1 - r2 * L.TL[1], zl = r2 - r3 * L.TL[2], tl = r3;
    int ty = 0, tz = 0, tt = 0;
    if (L.morton) {
        int pos = 0, maxb = std::max(L.tbits[1], std::max(L.tbits[2], L.tbits[3]));
        for (int b = 0; b < maxb; ++b) {
            if (b < L.tbits[1]) ty |= (int)((tau >> pos++) & 1u) << b;
            if (b < L.tbits[2]) tz |= (int)((tau >> pos++) & 1u) << b;
            if (b < L.tbits[3]) tt |= (int)((tau >> pos++) & 1u) << b;
        }
    } else {
        ty = tau % L.NT[1];
        tz = tau / L.NT[1] % L.NT[2];
        tt = tau / L.NT[1] / L.NT[2];
    }
    Coords c;
    c.y = ty * L.TL[1] + yl;
    c.z = tz * L.TL[2] + zl;
    c.t = tt * L.TL[3] + tl;
    c.x = 2 * xh + ((c.y + c.z + c.t + p) & 1);
    return c;
}

// slot of a link in the AoSoA-64 storage: class hs of S links, 64-link
// blocks of n slots, entry-major inside a block
static size_t plain_slot(uint32_t n, uint32_t S, int hs, uint32_t i) {
    return ((size_t)hs * (S / 64) + i / 64) * (64 * (size_t)n) + i % 64;
}

int main() {
    long bad_div = 0, n_div = 0, bad_nb = 0, n_nb = 0, bad_tc = 0, n_tc = 0, bad_sl = 0, n_sl = 0;
    std::mt19937_64 g(1);
    for (uint32_t d = 1; d < 70000; ++d) {
        const FastDiv f = make_fastdiv(d);
        const uint32_t ns[8] = {0u, 1u, d - 1, d, d + 1, 0xFFFFFFFFu, (uint32_t)g(), (uint32_t)(g() >> (g() % 64))};
        for (uint32_t x : ns) ++n_div, bad_div += fdiv(x, f) != x / d;
    }
    for (uint32_t d : {3u, 7u, 24u, 12345u, (1u << 31) - 1, (1u << 31) + 5, 0xFFFFFFFFu}) {
        const FastDiv f = make_fastdiv(d);
        for (int k = 0; k < 200000; ++k) {
            const uint32_t x = (uint32_t)g();
            ++n_div, bad_div += fdiv(x, f) != x / d;
        }
    }
    const int shapes[][4] = {{4, 4, 4, 4}, {8, 6, 10, 4}, {6, 12, 8, 10}, {16, 8, 8, 8}, {4, 6, 6, 6}, {2, 2, 2, 2}, {10, 1, 3, 2}};
    for (const auto& sh : shapes)
        for (int morton = 0; morton < 2; ++morton) {
            qcb_lattice l{};
            for (int k = 0; k < 4; ++k) l.L[k] = sh[k];
            const uint64_t V = (uint64_t)sh[0] * sh[1] * sh[2] * sh[3];
            l.V = (uint32_t)V;
            l.Vh = (uint32_t)(V / 2);
            l.S = (l.Vh + 63) & ~63u;
            l.esize = 4;
            default_tiling(sh, l.tile, &l.morton);
            if (!morton) l.morton = 0;
            const LatDev<float> L = make_desc<float>(&l);
            for (uint32_t q = 0; q < l.Vh; ++q)
                for (int p = 0; p < 2; ++p) {
                    const Coords a = tiled_coords(L, q, p), b = plain_tiled(L, q, p);
                    ++n_tc, bad_tc += a.x != b.x || a.y != b.y || a.z != b.z || a.t != b.t;
                    for (int d1 = 0; d1 < 4; ++d1)
                        for (int s1 = -1; s1 <= 1; s1 += 2) {
                            ++n_nb, bad_nb += nb_cb(L, a, d1, s1) != plain_nb(L, a, d1, s1, 0, 0);
                            for (int d2 = 0; d2 < 4; ++d2)
                                if (d2 != d1)
                                    for (int s2 = -1; s2 <= 1; s2 += 2)
                                        ++n_nb, bad_nb += nb_cb(L, a, d1, s1, d2, s2) != plain_nb(L, a, d1, s1, d2, s2);
                        }
                }
        }
    // aosoa_slot up to the largest lattice (4V < 2^31, so S <= 2^28) and the
    // largest slot count (9), every class of links (8) and of plane fields
    for (uint32_t n : {2u, 3u, 6u, 9u})
        for (uint32_t S : {64u, 192u, 1u << 20, (1u << 27) + 64u, 1u << 28})
            for (int hs = 0; hs < 12; ++hs) {
                const uint32_t is[6] = {0u, 1u, 63u, 64u, S - 1u, (uint32_t)(g() % S)};
                for (uint32_t i : is) ++n_sl, bad_sl += aosoa_slot(n, S, hs, i) != plain_slot(n, S, hs, i);
                for (int k = 0; k < 2000; ++k) {
                    const uint32_t i = (uint32_t)(g() % S);
                    ++n_sl, bad_sl += aosoa_slot(n, S, hs, i) != plain_slot(n, S, hs, i);
                }
            }
    // flow-stage block sites: at most 32; whole rows exactly when 30 or more
    // of the 32 site lanes can hold whole rows
    long bad_fb = 0;
    for (uint32_t lxh = 2; lxh <= 80; ++lxh) {
        const uint32_t fb = flow_block_sites(lxh), best = lxh <= 32 ? lxh * (32 / lxh) : 0;
        bad_fb += fb > 32 || (best >= 30 ? fb != best : fb != 32);
    }
    printf("fdiv %ld/%ld wrong, nb_cb %ld/%ld wrong, tiled_coords %ld/%ld wrong, aosoa_slot %ld/%ld wrong, "
           "flow_block_sites %ld/79 wrong\n",
           bad_div, n_div, bad_nb, n_nb, bad_tc, n_tc, bad_sl, n_sl, bad_fb);
    return bad_div || bad_nb || bad_tc || bad_sl || bad_fb;
}
