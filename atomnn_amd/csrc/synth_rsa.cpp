// Random sequential insertion of atoms into a periodic cell with a minimum separation.
// Host-only helper for the synthetic workload generator (atomnn_amd/synth.py); it is not
// part of the evaluation path.
//
// Candidates are uniform in fractional coordinates (mt19937_64 seeded by the caller).  A
// candidate is rejected when any accepted atom lies closer than dmin under the periodic
// metric (nearest of the 27 neighbouring images of the wrapped difference, which is exact
// for the mildly strained cells generated here).  A fractional bin grid keeps each test
// O(1).
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

extern "C" int synth_rsa(int n, const double* lv, double dmin, uint64_t seed, long long max_tries, double* pos_out) {
    // lv: lattice vectors a = lv[0..2], b = lv[3..5], c = lv[6..8]; R has them as columns.
    double R[3][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) R[r][c] = lv[3 * c + r];
    // perpendicular heights h_c = V / |a_p x a_q| decide the bin counts
    auto col = [&](int c, double* v) { v[0] = R[0][c]; v[1] = R[1][c]; v[2] = R[2][c]; };
    double a[3], b[3], cc[3];
    col(0, a);
    col(1, b);
    col(2, cc);
    auto crossn = [](const double* x, const double* y) {
        const double z0 = x[1] * y[2] - x[2] * y[1], z1 = x[2] * y[0] - x[0] * y[2], z2 = x[0] * y[1] - x[1] * y[0];
        return std::sqrt(z0 * z0 + z1 * z1 + z2 * z2);
    };
    const double vol = std::fabs(a[0] * (b[1] * cc[2] - b[2] * cc[1]) - a[1] * (b[0] * cc[2] - b[2] * cc[0]) +
                                 a[2] * (b[0] * cc[1] - b[1] * cc[0]));
    if (vol <= 0) return -1;
    const double h[3] = {vol / crossn(b, cc), vol / crossn(cc, a), vol / crossn(a, b)};
    int nb[3];
    for (int d = 0; d < 3; ++d) {
        nb[d] = (int)std::floor(h[d] / dmin);
        if (nb[d] < 1) nb[d] = 1;
        if (nb[d] > 256) nb[d] = 256;
    }
    const int nbins = nb[0] * nb[1] * nb[2];
    std::vector<std::vector<int>> bins(nbins);
    std::vector<double> frac(3 * (size_t)n);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double dmin2 = dmin * dmin;
    int placed = 0;
    long long tries = 0;
    while (placed < n) {
        if (++tries > max_tries) return -2;
        const double f[3] = {U(rng), U(rng), U(rng)};
        int bi[3];
        for (int d = 0; d < 3; ++d) {
            bi[d] = (int)(f[d] * nb[d]);
            if (bi[d] >= nb[d]) bi[d] = nb[d] - 1;
        }
        bool ok = true;
        // with fewer than 3 bins along an axis, scan every bin of that axis once
        const int lo0 = nb[0] < 3 ? 0 : -1, hi0 = nb[0] < 3 ? nb[0] - 1 : 1;
        const int lo1 = nb[1] < 3 ? 0 : -1, hi1 = nb[1] < 3 ? nb[1] - 1 : 1;
        const int lo2 = nb[2] < 3 ? 0 : -1, hi2 = nb[2] < 3 ? nb[2] - 1 : 1;
        for (int o0 = lo0; o0 <= hi0 && ok; ++o0)
            for (int o1 = lo1; o1 <= hi1 && ok; ++o1)
                for (int o2 = lo2; o2 <= hi2 && ok; ++o2) {
                    const int x = nb[0] < 3 ? o0 : ((bi[0] + o0) % nb[0] + nb[0]) % nb[0];
                    const int y = nb[1] < 3 ? o1 : ((bi[1] + o1) % nb[1] + nb[1]) % nb[1];
                    const int z = nb[2] < 3 ? o2 : ((bi[2] + o2) % nb[2] + nb[2]) % nb[2];
                    for (int q : bins[(x * nb[1] + y) * nb[2] + z]) {
                        double df[3];
                        for (int d = 0; d < 3; ++d) {
                            double t = f[d] - frac[3 * (size_t)q + d];
                            t -= std::nearbyint(t);
                            df[d] = t;
                        }
                        // nearest of the 27 images of the wrapped difference
                        double best = 1e300;
                        for (int s0 = -1; s0 <= 1; ++s0)
                            for (int s1 = -1; s1 <= 1; ++s1)
                                for (int s2 = -1; s2 <= 1; ++s2) {
                                    const double g0 = df[0] + s0, g1 = df[1] + s1, g2 = df[2] + s2;
                                    const double x0 = R[0][0] * g0 + R[0][1] * g1 + R[0][2] * g2;
                                    const double x1 = R[1][0] * g0 + R[1][1] * g1 + R[1][2] * g2;
                                    const double x2 = R[2][0] * g0 + R[2][1] * g1 + R[2][2] * g2;
                                    const double d2 = x0 * x0 + x1 * x1 + x2 * x2;
                                    if (d2 < best) best = d2;
                                }
                        if (best < dmin2) {
                            ok = false;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        for (int d = 0; d < 3; ++d) frac[3 * (size_t)placed + d] = f[d];
        bins[(bi[0] * nb[1] + bi[1]) * nb[2] + bi[2]].push_back(placed);
        ++placed;
    }
    for (int i = 0; i < n; ++i) {
        const double* f = &frac[3 * (size_t)i];
        for (int r = 0; r < 3; ++r) pos_out[3 * (size_t)i + r] = R[r][0] * f[0] + R[r][1] * f[1] + R[r][2] * f[2];
    }
    return 0;
}
