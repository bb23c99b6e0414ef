// Drives the oracle (test infrastructure) through every entry point of its C interface on small
// structures, for a build with AddressSanitizer and UndefinedBehaviorSanitizer (make sanitize):
// periodic cubic and sheared cells with several image shells, an open cluster, one atom, the
// min-image and all-images rules, energy training and force training with gradients.
// Usage: sanitize_oracle <potential.ann>...   Exit 0 when every call returned normally.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
const char* orc_last_error();
void* orc_load(const char* path);
void orc_free(void* m);
int orc_ntypes(void* m);
double orc_rc(void* m);
int orc_ninput(void* m, int t);
long long orc_neighbors(int n, const double* cell, const double* pos, const int* type, double rc, int* counts,
                        long long cap, int* index, int* image, int* is_min, double* r, double* d, int* shell);
int orc_symm(void* m, int n, const double* cell, const double* pos, const int* type, double* G, int ldg);
int orc_compute(void* m, int n, const double* cell, const double* pos, const int* type, int flags, double* energy,
                double* eatom, double* force);
void* orc_train_create(void* m, int nstruc, const int* natoms, const double* cells, const double* pos,
                       const int* type, const double* eref);
void orc_train_free(void* t);
int orc_train_nparams(void* t, int type);
int orc_train_get_params(void* t, double* x);
int orc_train_error(void* t, const double* x, int nbatch, const int* batch, int loss, double delta, double* err,
                    double* grad);
int orc_train_set_forces(void* t, const double* f);
int orc_train_error_force(void* t, const double* x, int nbatch, const int* batch, int loss, double delta, double beta,
                          int all_images, double* err, double* grad);
}

struct Case {
    int n;
    std::vector<double> cell;  // 9 values, all zero for an open structure
    std::vector<double> pos;
    std::vector<int> type;
};

static Case make_case(int n, double L, double shear, bool open, int ntypes, unsigned seed) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    Case c;
    c.n = n;
    c.cell.assign(9, 0.0);
    if (!open) {
        c.cell[0] = L;
        c.cell[4] = L;
        c.cell[8] = L;
        c.cell[3] = shear * L;  // b = (shear L, L, 0)
    }
    for (int i = 0; i < n; ++i) {
        for (int tries = 0; tries < 10000; ++tries) {
            const double f0 = u(rng), f1 = u(rng), f2 = u(rng);
            const double x = f0 * L + f1 * (open ? 0.0 : shear * L), y = f1 * L, z = f2 * L;
            bool ok = true;
            for (int j = 0; j < i && ok; ++j) {
                const double dx = x - c.pos[3 * j], dy = y - c.pos[3 * j + 1], dz = z - c.pos[3 * j + 2];
                ok = dx * dx + dy * dy + dz * dz > 1.5 * 1.5;
            }
            if (ok || tries == 9999) {
                c.pos.insert(c.pos.end(), {x, y, z});
                break;
            }
        }
        c.type.push_back(i % ntypes);
    }
    return c;
}

static int check(int rc, const char* what) {
    if (rc < 0) std::fprintf(stderr, "%s failed: %s\n", what, orc_last_error());
    return rc < 0 ? 1 : 0;
}

int main(int argc, char** argv) {
    int fails = 0;
    for (int a = 1; a < argc; ++a) {
        void* m = orc_load(argv[a]);
        if (!m) {
            std::fprintf(stderr, "load %s: %s\n", argv[a], orc_last_error());
            return 2;
        }
        const int T = orc_ntypes(m);
        const double rc = orc_rc(m);
        std::vector<Case> cases = {make_case(40, 9.0, 0.0, false, T, 1), make_case(12, 5.5, 0.25, false, T, 2),
                                   make_case(30, 8.0, 0.0, true, T, 3), make_case(1, 4.0, 0.0, false, T, 4),
                                   make_case(1, 4.0, 0.0, true, T, 5)};
        int ldg = 0;
        for (int t = 0; t < T; ++t) ldg = std::max(ldg, orc_ninput(m, t));
        for (const Case& c : cases) {
            const double* cell = c.cell[0] != 0.0 ? c.cell.data() : nullptr;
            std::vector<int> counts(c.n);
            int shell[3];
        const long long tot = orc_neighbors(c.n, cell, c.pos.data(), c.type.data(), rc, counts.data(), 0, nullptr,
                                            nullptr, nullptr, nullptr, nullptr, shell);
            fails += check((int)std::min<long long>(tot, 0), "neighbours (count)");
            std::vector<int> idx(tot), img(tot), mn(tot);
            std::vector<double> r(3 * tot), d(tot);
            fails += check((int)std::min<long long>(orc_neighbors(c.n, cell, c.pos.data(), c.type.data(), rc,
                                                                  counts.data(), tot, idx.data(), img.data(),
                                                                  mn.data(), r.data(), d.data(), nullptr), 0),
                           "neighbours");
            std::vector<double> G((size_t)c.n * ldg);
            fails += check(orc_symm(m, c.n, cell, c.pos.data(), c.type.data(), G.data(), ldg), "symm");
            for (int flags : {0, 1, 3, 7, 15}) {
                double E = 0;
                std::vector<double> ea(c.n), f(3 * c.n);
                fails += check(orc_compute(m, c.n, cell, c.pos.data(), c.type.data(), flags, &E, ea.data(), f.data()),
                               "compute");
                if (!std::isfinite(E)) ++fails;
            }
        }
        // training over the periodic cases
        std::vector<int> nat;
        std::vector<double> cells, pos, eref;
        std::vector<int> typ;
        for (const Case& c : cases) {
            nat.push_back(c.n);
            cells.insert(cells.end(), c.cell.begin(), c.cell.end());
            pos.insert(pos.end(), c.pos.begin(), c.pos.end());
            typ.insert(typ.end(), c.type.begin(), c.type.end());
            eref.push_back(-3.0 * c.n);
        }
        void* t = orc_train_create(m, (int)cases.size(), nat.data(), cells.data(), pos.data(), typ.data(), eref.data());
        if (!t) {
            std::fprintf(stderr, "train_create: %s\n", orc_last_error());
            ++fails;
        } else {
            int np = 0;
            for (int s = 0; s < T; ++s) np += orc_train_nparams(t, s);
            std::vector<double> x(np), g(np), f(pos.size(), 0.1);
            fails += check(orc_train_get_params(t, x.data()), "train params");
            std::vector<int> batch = {0, 1, 2, 3, 4};
            double err[3];
            for (int loss = 0; loss < 4; ++loss)
                fails += check(orc_train_error(t, x.data(), 5, batch.data(), loss, 0.05, err, g.data()), "train error");
            fails += check(orc_train_set_forces(t, f.data()), "train forces");
            for (int ai = 0; ai < 2; ++ai)
                fails += check(orc_train_error_force(t, x.data(), 5, batch.data(), 0, 1.0, 2.0, ai, err, g.data()),
                               "train error force");
            orc_train_free(t);
        }
        orc_free(m);
    }
    std::printf("sanitize_oracle: %d failed calls\n", fails);
    return fails == 0 ? 0 : 1;
}
